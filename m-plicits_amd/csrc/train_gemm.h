// Matrix-core GEMMs of the training step with fused epilogues (train_gemm.hip).
#pragma once

#include "common.h"

namespace mp {

// epilogues
enum {
  TG_ACT = 0,        // forward hidden layer: D, X = [sin; cos * tangent] unit-blocked, Xb row-blocked
  TG_OUT = 1,        // forward last sine layer + output layer: (f, grad f) partial sums per column range
  TG_OUT_BWD = 2,    // last sine layer recomputed + output-layer backward: Abar, Ab, (db, dw_out) partials
  TG_ACT_BWD = 3,    // backward through a hidden layer's activation: Abar, Ab, db partials
  TG_FIRST_BWD = 4,  // backward into layer 0 with its parameter gradients: (dW_0, db_0) partials
  TG_PART = 5,       // split-K weight-gradient partials
};

// C [M][N] = sum_k A(m, k) B(n, k), element type dtype (0 float32, 1 bfloat16,
// 2 float16), FP32 accumulation.  A is unit-blocked [K/KL][M][KL] and B
// blocked [K/KL][N][KL] (KL = 16 bytes / element size) for every epilogue but
// TG_PART, whose operands are the row-blocked copies (A [K/KL][M][KL] with
// K = the activation rows).  Rows 4p..4p+3 of M are the value and the three
// coordinate tangents of point p.
struct TgArgs {
  const void* A;
  const void* B;
  int64_t M;          // rows of C
  int N;              // columns of C
  int64_t K;
  const float* bias;  // TG_ACT, TG_OUT, TG_OUT_BWD: [N] added to the value rows
  const float* wl;    // TG_OUT, TG_OUT_BWD: output-layer weights [N]
  const float4* pt4;  // TG_OUT_BWD: seeds (dL/df, dL/dgrad f) per point; TG_FIRST_BWD: (x, y, z, 0) per point
  const float4* w0;   // TG_FIRST_BWD: (W_0[u][0..2], b_0[u]) per unit
  void* D;            // TG_ACT: out, TG_ACT_BWD: in; the activation's derivative factors
                      // (cos a, sin(a) ad_1..3), point-major [M/4][N][4], operand type
  void* X;            // TG_ACT, TG_OUT_BWD, TG_ACT_BWD: out, unit-blocked [N/KL][M][KL]
  void* Xb;           // same as X, row-blocked [M/KL][N][KL] (NULL: not written)
  float* part;        // partial sums, see tgemm_part_floats
  int splits;         // TG_PART: K chunks
};

int tgemm(int dtype, int epi, const TgArgs& a, hipStream_t s);

// Tile sizes (rows / columns of C per workgroup) and the rows / columns one
// wave covers, per element type.
int tgemm_tile_rows(int dtype);
int tgemm_tile_cols(int dtype);
inline int tgemm_wave_rows(int dtype) { return dtype == 0 ? 64 : 128; }
inline int tgemm_wave_cols(int dtype) { return dtype == 0 ? 64 : 128; }

// Floats of `part` an epilogue writes:
//   TG_OUT       [N / wave cols][M / 4][4]: (f, grad f) over one wave's columns
//   TG_OUT_BWD   [M / wave rows][2][N]: (db, dw_out) over one wave's rows
//   TG_ACT_BWD   [M / wave rows][N]: db
//   TG_FIRST_BWD [M / wave rows][4][N]: (dW_0[u][0..2], db_0[u])
//   TG_PART      [splits][M][N]
int64_t tgemm_part_floats(int dtype, int epi, int64_t M, int N, int splits);

// K chunks of the split-K weight gradient for rows x Wout x Win.
int tgemm_weight_grad_chunks(int dtype, int64_t rows, int Wout, int Win);

}  // namespace mp
