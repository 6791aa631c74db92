// Tile shapes and launch arguments of the emission projector and its
// transpose (projector.hip), shared with the batch planner (batch.hip), which
// fills the same argument records into device job tables.
#pragma once
#include "common.hpp"

namespace fl {

// ---------------------------------------------------------------------------
// Forward projector
// ---------------------------------------------------------------------------
constexpr int kFW = 64;                   // detector columns per workgroup (one per lane)
constexpr int kFH = 32;                   // detector rows per chunk
constexpr int kFWaves = 4;
constexpr int kFRows = kFH / kFWaves;     // rows per wave per chunk
constexpr int kFG = 8;                    // angles per workgroup (one accumulator each)
// One angle's footprint for a 64x32 tile: the sample points span at most
// hypot(31, 63) < 70.3 rows (and cols); with the 0.5 margins and the bilinear
// neighbour that is <= 74 rows and a pitch <= 75.
constexpr int kFMaxFootprint = 74 * 75;
// Footprint floats staged per reference: one angle always fits, and the union
// of a group of nearby angles usually does.  One reference: 32 KB, still four
// workgroups per CU (the registers allow four), and fewer unions split than
// with 24 KB (5.42 vs 5.68 ms at 256^3 x 360).  Two references: 24 KB each.
constexpr int fwd_lds_ref(int nref) { return nref == 1 ? 8192 : 6144; }
static_assert(kFMaxFootprint <= fwd_lds_ref(2) && kFMaxFootprint <= fwd_lds_ref(1), "one angle's footprint must fit");
static_assert(kFG * 4 * 64 <= fwd_lds_ref(2), "the final reduction reuses the staging area");
constexpr int kFRefChunk = 2;             // references staged together
constexpr size_t fwd_lds_bytes(int nref) { return sizeof(float) * (size_t)nref * fwd_lds_ref(nref); }
// Footprint columns the staging covers (two per lane); a union of several
// angles' footprints may not be wider.  One angle's is at most 75.
constexpr int kFStageCols = 2 * 64;
static_assert(75 <= kFStageCols, "one angle's footprint must fit the staged width");

enum FwdMode { kFwdWrite = 0, kFwdAccumulate = 1, kFwdRatio = 2, kFwdAccumulateRatio = 3 };

struct FwdArgs {
    const float *atten;
    const float *C;
    const float *em;
    const float *theta;
    const float *I;  // ratio modes
    float *out;      // (A, B, W): Pf or ratio
    int A, B, R, H, W;
    int r0, nref;
    int mode;
    int ntw, n_groups;
    int group_angles;  // grouped launches: angles per workgroup of this job (4 or 8)
    int out_sm;        // write out as (A, W, B) instead of (A, B, W)
    int out_pitch;     // slice-major output: slices per (a, w) row of out (0: B)
    int fold_em;       // ratio modes: write em[a] * ratio (the back-projector then takes em = 1)
};

// ---------------------------------------------------------------------------
// Back-projection (exact transpose) and the fused MLEM update
// ---------------------------------------------------------------------------
constexpr int kBI = 31;               // image rows per tile (base cells of its samples: 32 rows)
constexpr int kBJ = 63;               // image cols per tile (base cells: 64 cols)
constexpr int kBWaves = 4;            // private accumulators per workgroup
constexpr int kBPitch = kBJ + 2;      // 65: base-cell cols plus the right neighbour (odd)
// Accumulator cells of a tile BI rows high: its base cells plus one halo row
// below and one column to the right.
constexpr int bwd_acc_cells(int bi) { return (bi + 2) * kBPitch; }
constexpr int kBRefChunk = 2;
constexpr int kBSink = 136;           // sink cells: lane + corner offset <= 63 + 66
constexpr int kBItems = 32;           // items of 128 samples per enumeration window
constexpr int kBWin = 128 * kBItems;  // samples per window
constexpr int kBUnroll = 2;           // items splatted together
constexpr int kBRows = 48;            // detector rows per batch
static_assert(kBItems * 2 == 64, "one mask word per lane");
static_assert(kBItems % kBUnroll == 0, "items come in whole unroll groups");

// Dynamic LDS of the back-projector: four private accumulators (cell-major,
// nref values per cell), then one sink shared by the waves (it only ever holds
// discarded values, so their unordered writes there are harmless).  With the
// static tables this stays under 40 KB for nref = 1: four workgroups per CU.
constexpr size_t bwd_lds_bytes(int nref, int bi = kBI)
{
    return sizeof(float) * (size_t)nref * (kBWaves * bwd_acc_cells(bi) + kBSink);
}

enum BwdMode { kBwdWrite = 0, kBwdSens = 1, kBwdUpdate = 2, kBwdAccumulate = 3 };

struct BwdArgs {
    const float *atten;
    const float *ratio;  // (A,B,W) or null (= 1)
    const float *em;
    const float *theta;
    float *out;  // (R,B,H,W)
    int A, B, R, H, W;
    int r0, nref;
    int nti, ntj;
    int mode;
    const float *C;     // update: current estimate (R,B,H,W)
    const float *sens;  // update: sensitivity (R,B,H,W)
    float beta, delta;
    int tile_rows;  // grouped launches: image rows per tile of this job (15 or 31)
};

// ---------------------------------------------------------------------------
// Grouped (batched) launches: one persistent grid walks the tiles of many
// independent jobs.  Job j owns the logical tiles [begin[j], begin[j+1]);
// begin[0] = 0 and begin[n_jobs] = the total.
// ---------------------------------------------------------------------------
constexpr int kCopyTile = 4096;  // floats per tile of the grouped copy

struct CopyJob {
    const float *src;
    float *dst;
    int64_t n;
};

// Tiles of one job of the grouped launches.
inline uint32_t fwd_job_tiles(const FwdArgs &p) { return (uint32_t)(p.ntw * p.n_groups * p.B); }
inline uint32_t bwd_job_tiles(const BwdArgs &p) { return (uint32_t)(p.nti * p.ntj * p.B); }
inline uint32_t copy_job_tiles(const CopyJob &c) { return (uint32_t)((c.n + kCopyTile - 1) / kCopyTile); }

// Persistent grid sizes (resident workgroups on the current device).
int fwd_group_grid(int *grid);
int bwd_group_grid(int *grid);
int copy_group_grid(int *grid);
// Enqueue one grouped pass; jobs and begin are device tables.
int launch_fwd_group(const FwdArgs *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                     hipStream_t s);
int launch_bwd_group(const BwdArgs *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                     hipStream_t s);
int launch_copy_group(const CopyJob *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                      hipStream_t s);

// 4-neighbour Huber gradient of C at (i, j) in one (H, W) plane
// (FL/cuda_lib/atten_cuda_kernel.cu:593-613: neighbours up, down, left, right).
__device__ __forceinline__ float huber_at(const float *Cs, int i, int j, int H, int W, float delta)
{
    const float Cp = Cs[(int64_t)i * W + j];
    float g = 0.0f;
    const int di[4] = {-1, 1, 0, 0};
    const int dj[4] = {0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ii = i + di[k], jj = j + dj[k];
        if (ii < 0 || ii >= H || jj < 0 || jj >= W) continue;
        const float diff = Cp - Cs[(int64_t)ii * W + jj];
        g += fabsf(diff) <= delta ? diff : delta * copysignf(1.0f, diff);
    }
    return g;
}

// Slice-major back-projector (bwd_sm.hip): atten_sm (A, H, W, B), ratio_sm (A, W, B)
// or null, one reference per line (em (A)); out/C/sens (B, H, W); mode a BwdMode.
// spl: slices per lane (1, 2 or 4, dividing B), 0 for the default choice by B.
size_t sm_scratch_bytes(int64_t A, int64_t B, int64_t H, int64_t W, int spl = 0);
int launch_sm_backward(const float *atten_sm, const float *ratio_sm, const float *em, const float *theta,
                       float *out, int64_t A, int64_t B, int64_t H, int64_t W, int mode, const float *C,
                       const float *sens, float beta, float delta, void *scratch, size_t scratch_bytes,
                       hipStream_t stream);
// (N, B, L) -> (N, L, B)
int launch_sm_transpose(const float *src, float *dst, int64_t N, int64_t B, int64_t L, hipStream_t s);
// (N, B, L) -> slices 0..B-1 of (N, L, Bp), dst pointing at the first slice
int launch_sm_transpose_pitched(const float *src, float *dst, int64_t N, int64_t B, int64_t L, int64_t Bp,
                                hipStream_t s);
// ratio[(a * W + w) * Bp + b] = em[a] for b < B (the back-projection of em: the sensitivity)
int launch_sm_fill_em(float *ratio, const float *em, int64_t A, int64_t W, int64_t B, int64_t Bp, hipStream_t s);

// Grouped slice-major back-projection over many problems (batched MLEM): every
// problem ("pack") has its own atten_sm, ratio (A, W, B), em, theta and scratch;
// the tables are built on the host once and replayed.  Ratio is always given.
struct SmArgs;
struct SmFin;
struct SmPack {
    const float *atten_sm, *ratio, *em, *theta;
    int64_t A, B, H, W;
    int spl;        // slices per lane: 1, 2 or 4, dividing B
    void *scratch;  // sm_scratch_bytes(A, B, H, W, spl)
};
struct SmFinSpec {  // one finalize variant of one pack
    float *out;
    int mode;  // BwdMode
    const float *C, *sens;
    float beta, delta;
};
constexpr int kSmGroupVariants = 3;
struct SmGroup {
    struct Class {
        int spl = 0, n = 0;
        size_t jobs = 0, b_fused = 0, b_red = 0;  // byte offsets in the table image
        uint32_t t_fused = 0, t_red = 0;           // workgroups per launch
    } cls[3];
    int n = 0, n_var = 0;
    size_t fin[kSmGroupVariants] = {}, b_fin = 0;
    uint32_t t_fin = 0;
    const char *dev = nullptr;  // device address of the table image
};
size_t sm_group_table_bytes(int n);
// Writes the tables of n packs into host[0, sm_group_table_bytes(n)); fin holds
// n_var finalize variants, fin[v * n + i] for pack i.  out->dev is set by the caller.
int sm_group_build(const SmPack *packs, int n, const SmFinSpec *fin, int n_var, char *host, SmGroup *out);
// Back-projection of every pack, then finalize variant `variant`.
int launch_sm_group(const SmGroup &g, int variant, hipStream_t s);

}  // namespace fl
