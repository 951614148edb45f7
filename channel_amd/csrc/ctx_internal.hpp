// ctx_internal.hpp -- the state behind a chan_ctx handle, shared by the
// translation units that implement its entry points (api.hip, k_les.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/channel_hip.h"
#include "chan_internal.hpp"

namespace chan {
struct LesState;
// frees the closure buffers of a context (k_les.hip)
void les_free(LesState *s);
}  // namespace chan

struct chan_ctx {
  chan_params prm;
  chan::HostSetup hs;
  chan::Geom g;
  hipStream_t st = nullptr;
  // device state
  chan::cx<double> *V = nullptr, *rhs = nullptr, *col00 = nullptr;
  void *old = nullptr;  // explicit terms: complex FP64, or FP32 in the split mode
  double2 *L = nullptr;
  // transform intermediates and twiddles: complex FP64, or complex FP32 in
  // the split mode
  bool split = false;
  void *A = nullptr, *B = nullptr, *P = nullptr, *twz = nullptr, *twx = nullptr;
  double *der = nullptr, *d0fac = nullptr, *y = nullptr, *dy = nullptr;
  double *stat = nullptr;  // statistics: [3][ny+1] mean in, [7][ny+1] profiles out
  // body force: static field [3][Ny][ps] and feedback gains [3][Ny][Nz+1]
  chan::cx<double> *F0 = nullptr;
  double *gain = nullptr;
  int src[3] = {-1, -1, -1};
  // the force of the last buildrhs, kept on request: [3][Ny][ps]; kept_state
  // 0 = nothing kept yet, 1 = Fkeep holds it, 2 = that buildrhs had no force
  chan::cx<double> *Fkeep = nullptr;
  int kept_state = 0;
  chan::DevScalars *sc = nullptr;
  // host scalars
  double time = 0, deltat = 0;
  bool field_set = false;
  // timing
  bool timing = false;
  struct Rec {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Rec> pend;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;
  std::vector<long long> counts;
  std::vector<double> ms;
  // the learned closure's device buffers, made on first use (k_les.hip)
  chan::LesState *les = nullptr;
};

