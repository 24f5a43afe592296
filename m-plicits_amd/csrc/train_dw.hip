// Weight gradients of the hidden layers of the training step,
//   dW [Wout][Win] = sum over rows of A[row][i] X[row][j]
// (rows = points x (1 + 3 tangents), K = 2^20 at the 2^18-point batch), from
// the K-blocked 16-bit copies the elementwise stages write beside their
// row-major outputs: element (row r, unit u) at ((r / 8) * W + u) * 8 + r % 8,
// so the eight reduction elements an MFMA lane needs for one unit are 16
// contiguous bytes.  (The reference computes these with autograd's GEMMs,
// experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252.)
//
// Split-K: every workgroup (four waves, one per SIMD, each a 128 x 128 tile
// with its accumulators in AGPRs) reduces one K chunk into a 256 x 256 output
// tile and writes FP32 partial sums part[chunk][Wout][Win]; the caller adds
// the chunks.  The tiles of one chunk go to workgroups on the same XCD
// (workgroup ids b, b + 8, ... share one), so the chunk's operand rows are
// fetched from HBM into that XCD's L2 once.  The operands stream through a
// DW_RING-deep register ring (8 fragments = 8 KB per wave per K step) so that
// enough loads are in flight to cover HBM latency; loads past the chunk's end
// re-read its last step instead of branching.  Operand pointers are not
// __restrict__ and every ring load ends in a compiler memory barrier, so the
// loads stay DW_RING steps ahead of their MFMAs.
#include "common.h"
#include "mlp_core.h"

namespace mp {

constexpr int DW_RING = 4;  // K steps in flight per wave; chunks hold multiples of it

template <class S>
__global__ __launch_bounds__(256) void k_train_dw(const typename S::Frag* Ab, const typename S::Frag* Xb,
                                                  int64_t nsteps,
                                                  int Wout, int Win, int chunks, int64_t steps_per_chunk,
                                                  float* __restrict__ part) {
  typedef typename S::Frag Frag;
  typedef typename S::Acc Acc;
  const int tiles_j = Win / 256, tiles = (Wout / 256) * tiles_j;
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int tile = slot % tiles;
  const int chunk = xcd + 8 * (slot / tiles);
  if (chunk >= chunks) return;  // workgroup-uniform
  const int ti = tile / tiles_j, tj = tile % tiles_j;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, col = lane & 31;
  const int i0 = ti * 256 + (wave >> 1) * 128, j0 = tj * 256 + (wave & 1) * 128;
  Acc acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[a][b][q] = 0.f;
  const int64_t s0 = (int64_t)chunk * steps_per_chunk;
  const int64_t s1 = s0 + steps_per_chunk < nsteps ? s0 + steps_per_chunk : nsteps;
  float* out = part + (int64_t)chunk * Wout * Win;
  if (s0 < s1) {
    // one K step = 16 rows = blocks 2s (lanes 0-31) and 2s + 1 (lanes 32-63)
    const Frag* pa = Ab + (2 * s0 + g) * Wout + i0 + col;
    const Frag* pb = Xb + (2 * s0 + g) * Win + j0 + col;
    const Frag* const pa_last = Ab + (2 * (s1 - 1) + g) * Wout + i0 + col;
    const Frag* const pb_last = Xb + (2 * (s1 - 1) + g) * Win + j0 + col;
    const int64_t da = 2 * (int64_t)Wout, db = 2 * (int64_t)Win;
    Frag ra[DW_RING][4], rb[DW_RING][4];
    auto load = [&](int k) {
#pragma unroll
      for (int r = 0; r < 4; ++r) ra[k][r] = pa[32 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) rb[k][c] = pb[32 * c];
      pa = pa + da <= pa_last ? pa + da : pa_last;
      pb = pb + db <= pb_last ? pb + db : pb_last;
      // keeps the compiler from sinking the ring's loads to their first use
      asm volatile("" ::: "memory");
    };
#pragma unroll
    for (int k = 0; k < DW_RING; ++k) load(k);
    for (int64_t s = s0; s < s1; s += DW_RING) {
#pragma unroll
      for (int k = 0; k < DW_RING; ++k) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = S::mma(ra[k][r], rb[k][c], acc[r][c]);
        load(k);
      }
    }
  }
  // C layout of a 32 x 32 block: lane (column col, group g), register q holds
  // row 8 (q / 4) + 4 g + q % 4
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = i0 + 32 * r + 8 * (q >> 2) + 4 * g + (q & 3);
        out[(int64_t)row * Win + j0 + 32 * c + col] = acc[r][c][q];
      }
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_train_dw_chunks(int64_t rows, int Wout, int Win) {
  if (rows <= 0 || Wout < 256 || Win < 256) return 0;
  const int tiles = (Wout / 256) * (Win / 256);
  const int64_t nsteps = rows / 16;
  int chunks = (256 + tiles - 1) / tiles;  // about one workgroup per CU
  chunks = (chunks + 7) / 8 * 8;
  if (chunks > nsteps) chunks = (int)nsteps;
  return chunks < 1 ? 1 : chunks;
}

int mp_train_dw(const void* Ab, const void* Xb, int64_t rows, int Wout, int Win, int dtype, float* part,
                void* stream) {
  MP_REQUIRE(Ab && Xb && part, MP_EINVAL, "mp_train_dw: NULL argument");
  MP_REQUIRE(dtype == 1 || dtype == 2, MP_EINVAL, "mp_train_dw: operands must be bfloat16 (1) or float16 (2)");
  MP_REQUIRE(rows > 0 && rows % (16 * DW_RING) == 0, MP_EINVAL,
             "mp_train_dw: rows must be a positive multiple of %d", 16 * DW_RING);
  MP_REQUIRE(Wout % 256 == 0 && Win % 256 == 0, MP_EUNSUPPORTED,
             "mp_train_dw: widths %d x %d (needs multiples of 256)", Wout, Win);
  const int chunks = mp_train_dw_chunks(rows, Wout, Win);
  const int64_t nsteps = rows / 16;
  const int64_t per = ((nsteps + chunks - 1) / chunks + DW_RING - 1) / DW_RING * DW_RING;
  const int tiles = (Wout / 256) * (Win / 256);
  const unsigned grid = (unsigned)(8 * tiles * ((chunks + 7) / 8));
  hipStream_t s = as_stream(stream);
  if (dtype == 1)
    hipLaunchKernelGGL(k_train_dw<S32x32x16_bf16>, dim3(grid), dim3(256), 0, s,
                       reinterpret_cast<const bf16x8*>(Ab), reinterpret_cast<const bf16x8*>(Xb), nsteps, Wout, Win,
                       chunks, per, part);
  else
    hipLaunchKernelGGL(k_train_dw<S32x32x16_f16>, dim3(grid), dim3(256), 0, s,
                       reinterpret_cast<const f16x8*>(Ab), reinterpret_cast<const f16x8*>(Xb), nsteps, Wout, Win,
                       chunks, per, part);
  MP_LAUNCHED("k_train_dw");
  return MP_OK;
}

}  // extern "C"
