// Driver of the library's host set-up (channel_amd/csrc/setup.cpp) for the
// sanitizer build of tests/test_host_sanitize.py: g++ with AddressSanitizer
// and UndefinedBehaviorSanitizer, every report fatal.
//
//   sanitize_setup OUT nx ny nz alfa0 beta0 a ymin ymax
//
// writes to OUT, as raw doubles: nxd, nzd, y [ny+3], der [(ny-1)][4][5],
// bnd [10][5], bc [8][5], d0fac [(ny+1)][5] (the outputs of
// chan_setup_host, in the same order), then fftfit(i) for i = 0..4096.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "setup.hpp"

int main(int argc, char **argv) {
  if (argc != 10) {
    std::fprintf(stderr, "usage: %s OUT nx ny nz alfa0 beta0 a ymin ymax\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
  const double alfa0 = std::atof(argv[5]), beta0 = std::atof(argv[6]), a = std::atof(argv[7]);
  const double ymin = std::atof(argv[8]), ymax = std::atof(argv[9]);
  chan::HostSetup h;
  chan::setup_build(h, nx, ny, nz, alfa0, beta0, a, ymin, ymax);
  std::vector<double> out;
  out.push_back(h.nxd);
  out.push_back(h.nzd);
  out.insert(out.end(), h.y, h.y + (ny + 3));
  out.insert(out.end(), h.der, h.der + 20 * (ny - 1));
  out.insert(out.end(), &h.bnd[0][0], &h.bnd[0][0] + 50);
  out.insert(out.end(), &h.bc[0][0], &h.bc[0][0] + 40);
  out.insert(out.end(), h.d0fac, h.d0fac + 5 * (ny + 1));
  for (int i = 0; i <= 4096; i++) out.push_back(chan::fftfit(i));
  chan::setup_free(h);
  FILE *f = std::fopen(argv[1], "wb");
  if (!f) return 3;
  const size_t n = std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  return n == out.size() ? 0 : 4;
}
