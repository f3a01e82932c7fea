// Checkpoint buffers (src/IO/io_buffer.c:26-81): one row per cell, in the
// reference's cell order (layer k, annulus i, cell j; ghosts included), of
// NUM_Q + 3 doubles: tiph, the annulus centre r, the layer centre z, the
// primitives.  Packed and unpacked on the GPU so a checkpoint moves as one
// contiguous copy.  One workgroup per annulus of the disk.
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace {

__global__ __launch_bounds__(TILE) void k_ckpt_pack(DevMesh m, int disk, double *out) {
  const DiskParams &P = m.dp[disk];
  const AnnInfo A = m.ann[P.ann_base + blockIdx.x];
  const int nc = P.NQ + 3;
  const long long z0 = m.ann[P.ann_base].zbase;
  const double r = 0.5 * (A.rm + A.rp), zc = 0.5 * (A.zm + A.zp);
  for (int j = threadIdx.x; j < A.N; j += TILE) {
    const long long c = A.zbase + j;
    double *o = out + (c - z0) * nc;
    o[0] = m.tiph[c];
    o[1] = r;
    o[2] = zc;
    for (int q = 0; q < P.NQ; ++q) o[3 + q] = m.prim[q * m.Ntot + c];
  }
}

// io_readbuf (io_buffer.c:57-81): tiph and the primitives; v_z = 0 in 2D
__global__ __launch_bounds__(TILE) void k_ckpt_unpack(DevMesh m, int disk, const double *in, double *tiph) {
  const DiskParams &P = m.dp[disk];
  const AnnInfo A = m.ann[P.ann_base + blockIdx.x];
  const int nc = P.NQ + 3;
  const long long z0 = m.ann[P.ann_base].zbase;
  for (int j = threadIdx.x; j < A.N; j += TILE) {
    const long long c = A.zbase + j;
    const double *v = in + (c - z0) * nc;
    tiph[c] = v[0];
    for (int q = 0; q < P.NQ; ++q) m.prim[q * m.Ntot + c] = v[3 + q];
    if (P.NglobalZ == 1) m.prim[UZZ * m.Ntot + c] = 0.0;
  }
}

}  // namespace

void launch_ckpt_pack(const DevMesh &m, int disk, int nann, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_ckpt_pack, dim3(nann), dim3(TILE), 0, st, m, disk, out);
}
void launch_ckpt_unpack(const DevMesh &m, int disk, int nann, const double *in, double *tiph, hipStream_t st) {
  hipLaunchKernelGGL(k_ckpt_unpack, dim3(nann), dim3(TILE), 0, st, m, disk, in, tiph);
}
