"""CPU oracle for the particle-filter step and the batched WLS.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (gnss_gpu_amd/) imports this
package; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it, and only as the checker or the timed CPU baseline.

Contents
  philox.py   NumPy twin of the library's Philox4x32-10 streams and uniforms.
  detmath.py  bit-exact twins of the deterministic pieces of the GPU
              resampler (binary64/binary32 exp, the fixed-order block scans,
              the tile/decision/CDF pipeline, Megopolis).
  pf.py       binary64 restatement of the reference particle filter
              (src/particle_filter/pf_device.cu) driven by the same random
              numbers as the GPU.
  wls.py      binary64 restatement of src/positioning/wls.cu.
  cpu_pf.c    C restatement used as the timed CPU baseline (OpenMP).

Each function cites the reference file:line it follows.  pinning.py lists,
for every function, the check that pins it; what only an import of the
reference could pin (its cuRAND streams, thrust's scan order) is listed there
as parity unpinned.
"""
