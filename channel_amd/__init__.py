"""channel_amd -- MI355X-native time step of the channel DNS.

The hot path (the RK3 step: dealiased x/z FFTs, convolution products,
compact d/dy stencils, batched pentadiagonal v-eta solves, on-device
transposes) lives in the HIP C-ABI library ``libchannel_hip.so`` built from
``channel_amd/csrc``.  ``channel_amd.lib`` loads it; ``channel_amd.dns`` holds
the reference's file formats (dns.in, Dati.cart.out, Runtimedata);
``channel_amd.solver`` mirrors the reference's per-step interface
(buildrhs / linsolve / outstats / the time loop) on top of the library.
"""
