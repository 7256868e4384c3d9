/*
 * lbm3d.h -- C ABI of liblbm3d.so, the MI355X (gfx950) single-phase sparse
 * D3Q19 MRT lattice-Boltzmann path.
 *
 * The reference exposes this path as the Python class LB3D_Solver_Single_Phase
 * (Single_phase/LBM_3D_SinglePhase_Solver.py, "SP" below) whose Taichi
 * kernels run on the Taichi runtime.  These entry points replace those
 * kernels; the Python host class in taichi_lbm3d_amd/solver.py keeps the
 * reference's surface and calls them through ctypes.
 *
 * Conventions shared by every entry point:
 *   - All pointers except `p` are device pointers owned by the caller; the
 *     library never allocates or frees memory and keeps no pointer between calls.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Work is only
 *     enqueued; nothing synchronises unless stated.
 *   - Return 0 on success, a negative LBM3D_E* code otherwise;
 *     lbm3d_last_error() then gives a message (per thread).
 *   - Fluid cells (solid value 0) are numbered 0 .. n_cells-1 in the list
 *     order that p->band_rows sets (lbm3d_compact_fill): every grid row's
 *     cells get consecutive ids, k fastest; band_rows = 0 is C order of
 *     (i, j, k).
 *   - Populations are stored shifted, g_q = f_q - w_q, in SoA order
 *     g[q * S + cell], q = 0..18 in the reference's direction order
 *     (SP:183-187), S = lbm3d_slot_stride(n_cells) (n_cells rounded up to a
 *     multiple of 64), in the storage type `dtype` (LBM3D_F32/F16/BF16/F64),
 *     so g holds 19 * S elements; arithmetic is FP32 (FP64 for LBM3D_F64).
 *   - face_u (the pressure-face state) and the lbm3d_macros outputs are FP32,
 *     or FP64 when dtype is LBM3D_F64.
 *   - Between steps g is in the AA pattern's layout for the step count t:
 *     even t, population q of a cell in its own slot q; odd t, in slot
 *     opp(q) of the upstream neighbour (or in the cell's own slot q where
 *     that neighbour is not fluid).
 */
#ifndef LBM3D_H
#define LBM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM3D_ABI_VERSION 14

/* storage types for populations; F32, F16 and BF16 compute in FP32, F64 in FP64 */
#define LBM3D_F32 0
#define LBM3D_F16 1
#define LBM3D_BF16 2
#define LBM3D_F64 3

/* lbm3d_params.flags: F16/BF16 population stores of the step kernels round
 * stochastically (unbiased: each stored value equals the exact one in
 * expectation) instead of to nearest, with random bits that depend only on
 * the step t, the writing cell's id within its own grid (a batch sample's ids
 * count from its cell_begin) and the store's place in the cell's fixed store
 * order.  A run is reproducible, and a batch sample gets the bits of the same
 * grid run alone.  The step's seed is computed on the host from t, so a
 * captured launch replays the bits of the step it was captured at: replaying
 * a graph of steps with this flag repeats them (BatchSolver refuses it).  It
 * keeps per-step increments far below the 16-bit spacing, e.g. of a weak body
 * force, from rounding away.  No effect on F32/F64 storage. */
#define LBM3D_FLAG_STOCHASTIC_ROUNDING 1
/* lbm3d_params.flags: the row table given to the step functions is followed,
 * at rows + lbm3d_row_words(p), by the lbm3d_tile_words(p) tile descriptors
 * that lbm3d_build_tiles wrote for the same cell list.  Odd steps then take
 * each block's rows from its descriptor and each cell's place from the row
 * words instead of reading the cell list first.  Results are identical.
 * Batches: the descriptors follow the batch row table (lbm3d_batch_sizes.
 * row_words), one per 256-cell block of the batch's cell range; sample s's
 * start at row_words + cell_begin / 256, written by lbm3d_build_tiles with the
 * sample's own params and cells.  Batch odd steps launched with 256-thread
 * blocks and no persistent grid use them. */
#define LBM3D_FLAG_TILES 2

/* error codes */
#define LBM3D_OK 0
#define LBM3D_EINVAL (-1) /* bad argument (null pointer, size, mode) */
#define LBM3D_EHIP (-2)   /* a HIP runtime call failed */
#define LBM3D_ESIZE (-3)  /* a size exceeds what the kernels index */

/* Face order everywhere: 0 x-left, 1 x-right, 2 y-left, 3 y-right, 4 z-left,
 * 5 z-right -- the order in which SP's Boundary_condition applies them
 * (SP:274-370), so a node on several faces belongs to the last one. */

/* Frozen solver parameters: what SP bakes into its kernels at
 * init_simulation() and first compilation (SP:118-149, SP:230, SP:274-365). */
/* Real parameters are doubles, as SP computes them in Python; a kernel rounds
 * them once to its arithmetic type (FP32 for F32/F16/BF16 storage, as SP's
 * f32 fields hold them, SP:54, :131). */
typedef struct lbm3d_params {
  int32_t nx, ny, nz;      /* grid extent (SP:11) */
  int32_t n_cells;         /* fluid cells after compaction */
  int32_t force_on;        /* force_flag (SP:137-140) */
  int32_t bc_mode[6];      /* 0 periodic, 1 fixed pressure, 2 fixed velocity (SP:22) */
  int32_t face_u_off[6];   /* element offset of each pressure face's state in face_u, -1 if unused */
  int32_t flags;           /* LBM3D_FLAG_* bits (0: none) */
  double s_v, s_o;         /* relaxation rates s_v and s_other (SP:127-129) */
  double fx, fy, fz;       /* body force (SP:17, SP:134-136) */
  double guo_a, guo_b;     /* force-term scalings: SP uses 1/3 and 1/9 (SP:236) */
  double rho_bc[6];        /* rho_bcxl ... (SP:23-28) */
  double u_bc[6][3];       /* vx_bcxl, vy_bcxl, vz_bcxl ... (SP:23-28) */
  int32_t band_rows;       /* order of the fluid list (lbm3d_compact_fill): bands of this many
                              rows j; 0 or >= ny: C order */
} lbm3d_params;

/* One 32-cell word of a grid row (i, j): bit b of `mask` set when cell
 * (i, j, 32w + b) is fluid; `base` = compact id of the first fluid cell of
 * the word; `left` / `right` = compact id of the cell just before 32w and
 * just after 32w+31 along k (with the periodic wrap of SP:247-257 at k = 0
 * and k = nz-1), or -1 when that cell is not fluid.
 * rows[(i * ny + j) * W + w], W = (nz + 31) / 32. */
typedef struct lbm3d_row_word {
  uint32_t mask;
  int32_t base;
  int32_t left;
  int32_t right;
} lbm3d_row_word;

int lbm3d_abi_version(void);
const char *lbm3d_last_error(void);

/* ---- sizes ---------------------------------------------------------------- */

/* Elements per population slot for n_cells fluid cells (a multiple of 64). */
int64_t lbm3d_slot_stride(int32_t n_cells);
/* Entries of the row table: nx * ny * ((nz + 31) / 32). */
int64_t lbm3d_row_words(const lbm3d_params *p);
/* Tile descriptors (lbm3d_row_word-sized entries) for p->n_cells fluid cells:
 * one per block of cells of the step kernels; 0 when n_cells is 0. */
int64_t lbm3d_tile_words(const lbm3d_params *p);
/* uint32 words of face_bits: 2 bits per fluid cell, (n_cells + 15) / 16. */
int64_t lbm3d_face_bit_words(int32_t n_cells);
/* Nodes in the plane of face f (ny*nz, nx*nz or nx*ny).  The state of a
 * pressure face f is 3 * lbm3d_face_plane(p, f) floats at face_u +
 * face_u_off[f]: component c of the node with in-plane index
 * (x faces j*nz + k, y faces i*nz + k, z faces i*ny + j) at c * plane + index. */
int64_t lbm3d_face_plane(const lbm3d_params *p, int32_t face);

/* ---- geometry: the compacted fluid list (replaces SP's pointer/dense SNode
 * tree, SP:41-44, and the fluid test at SP:225, :262, :376) ---------------- */

/* Workspace bytes that lbm3d_compact_count / lbm3d_compact_fill need for a
 * grid of n_dense = nx*ny*nz cells. */
size_t lbm3d_compact_workspace_bytes(int64_t n_dense);

/* Pass 1: count the fluid cells of `solid` (the int8 C-order (nx,ny,nz) field
 * of SP:57; a cell is fluid iff its value is 0) into count[0] (device int32). */
int lbm3d_compact_count(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *count,
                        void *workspace, size_t workspace_bytes);

/* Pass 2 (after pass 1 with the same workspace; p->n_cells = the count):
 * cells[n_cells] = C-order dense index of each fluid cell, and, when d2c is
 * not NULL, d2c[nx*ny*nz] = compact id or -1.  Compact ids follow the list
 * order of p->band_rows = R: the grid rows (i, j) are cut into bands of R
 * consecutive j (the last band may be narrower), and the fluid cells are
 * numbered band after band, inside a band plane after plane (i), then row
 * after row (j), k fastest.  R = 0 or R >= ny is C order (k fastest, then j,
 * then i).  Every row's fluid cells get consecutive ids in k order, which is
 * what the row table encodes, so every other function takes either order;
 * the odd step's neighbours in planes i +- 1 are R rows of cells away in the
 * list instead of ny rows. */
int lbm3d_compact_fill(void *stream, const lbm3d_params *p, const int8_t *solid, int32_t *d2c,
                       int32_t *cells, void *workspace, size_t workspace_bytes);

/* Row table (see lbm3d_row_word) from solid and the d2c map of pass 2. */
int lbm3d_build_rows(void *stream, const lbm3d_params *p, const int8_t *solid, const int32_t *d2c,
                     lbm3d_row_word *rows);
/* The same with every compact id (base, left, right) shifted by id_offset:
 * the row table of one sample of a batch, whose cells start at id_offset. */
int lbm3d_build_rows_at(void *stream, const lbm3d_params *p, const int8_t *solid, const int32_t *d2c,
                        int32_t id_offset, lbm3d_row_word *rows);

/* Tile descriptors (LBM3D_FLAG_TILES) from the fluid list of pass 2 (in the
 * order of p->band_rows), one per block of cells: {i, j, J, 0} when the
 * block's cells lie in rows (i, j .. j+J) of one plane's segment of the list;
 * {i, j, JA | JB << 8 | 0x10000, iB | jB << 16} when they lie in rows
 * (i, j .. j+JA) and then in the next segment's rows (iB, jB .. jB+JB), the
 * next plane of the band or the first plane of the next band; in both cases
 * only when their row window fits the odd step's LDS copy.  Otherwise
 * {i, j, 0xFFFFFFFF, 0}.  No reference counterpart: an index of the row table
 * for the odd step. */
int lbm3d_build_tiles(void *stream, const lbm3d_params *p, const int32_t *cells, lbm3d_row_word *tiles);

/* face_bits: per fluid cell, bit 2*(x%16) of word x/16 set when the cell
 * belongs to a face in mode 1 or 2 (owner = last such face in face order),
 * bit 2*(x%16)+1 set when the owner's inward neighbour (index 1 or n-2 along
 * the face normal) has solid value > 0 (SP:278, :294, ...). */
int lbm3d_build_face_bits(void *stream, const lbm3d_params *p, const int8_t *solid,
                          const int32_t *cells, uint32_t *face_bits);

/* Explicit neighbour table, SoA nbr[(q-1) * n_cells + cell] for q = 1..18:
 * the compact id of the cell at (i,j,k)+e_q under the unconditional periodic
 * wrap of SP's periodic_index (SP:247-257), or -1 when that cell is not
 * fluid, i.e. when streaming1 bounces f_q back into the cell's own opposite
 * slot (SP:265-268).  The step kernel computes the same ids from `rows` on
 * the fly, with the same device function; this table is for inspection and
 * tests. */
int lbm3d_build_neighbors(void *stream, const lbm3d_params *p, const int32_t *cells,
                          const lbm3d_row_word *rows, int32_t *nbr);

/* ---- solver ------------------------------------------------------------- */

/* init (SP:160-170): g = feq(rho0, u0) - w on every fluid cell, and the
 * pressure-face state = u0 at the face nodes.  rho0 [n_cells] and u0 (SoA
 * [3][n_cells], FP32) may be NULL for the reference's rest state rho = 1, u = 0. */
int lbm3d_init(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells,
               void *face_u, const float *rho0, const float *u0);

/* One time step: the reference's colission -> streaming1 ->
 * Boundary_condition -> streaming3 (SP:477-481) as one fused kernel over the
 * fluid list, in place with the AA access pattern.  `t` is the number of
 * steps already taken; its parity selects the access pattern, and t == 0 is
 * the step that starts from init's state (macros rho0, u0 without the
 * half-force shift, face populations as initialised). */
int lbm3d_step(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells,
               const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u, int64_t t);

/* `nsteps` consecutive lbm3d_step calls from step t0, enqueued back to back. */
int lbm3d_run(void *stream, const lbm3d_params *p, int dtype, void *g, const int32_t *cells,
              const lbm3d_row_word *rows, const uint32_t *face_bits, void *face_u, int64_t t0,
              int32_t nsteps);

/* Macroscopic fields after t steps, what SP's streaming3 leaves in rho and v
 * (SP:376-388): rho[n_cells] and v (SoA [3][n_cells]), FP32 (FP64 for
 * LBM3D_F64).  lbm3d_scatter_dense and lbm3d_reduce take the FP32 form. */
int lbm3d_macros(void *stream, const lbm3d_params *p, int dtype, const void *g, const int32_t *cells,
                 const lbm3d_row_word *rows, const uint32_t *face_bits, const void *face_u, int64_t t,
                 void *rho, void *v);

/* Scatter compact rho/v into dense C-order arrays of extent (ox,oy,oz) >=
 * (nx,ny,nz): rho_dense[ox*oy*oz], v_dense[ox*oy*oz][3] (the to_numpy layout).
 * v is SoA with component c of cell x at v[c * v_stride + x] (v_stride 0 means
 * n_cells).  Cells that are not fluid are left untouched. */
int lbm3d_scatter_dense(void *stream, const lbm3d_params *p, const int32_t *cells,
                        const float *rho, const float *v, int64_t v_stride, int32_t ox, int32_t oy,
                        int32_t oz, float *rho_dense, float *v_dense);

/* No reference counterpart (SP stores FP32): the rounding the step kernels
 * apply when they store a population in 16 bits, on n FP32 values in[] ->
 * out[] (the 16-bit patterns), dtype LBM3D_F16 or LBM3D_BF16.  stochastic 0:
 * round to nearest even.  Otherwise the rounding of
 * LBM3D_FLAG_STOCHASTIC_ROUNDING with the 16 random bits rnd[i] & 0xFFFF: BF16
 * adds them below its last place and truncates; FP16 adds the top 13 of them
 * below its last place for normal results (their top 13 + j for results in its
 * subnormal range, j the extra dropped bits, zero-filled below bit 16) and
 * truncates, and below 2^-24 stores 0 or +-2^-24 with the matching
 * probability.  Each result is one of the two values around in[i], the upper
 * with probability equal to the dropped fraction: over all 65,536 values of
 * the random bits the results average to in[i] exactly for BF16 and for FP16
 * normal results, and for FP16 results below 2^-14 (its subnormal range) to
 * in[i] less 1e-5 to 1e-4 of a step (gfx950's carry there comes up slightly
 * short).  For tests and for converting fields to the storage formats. */
int lbm3d_round16(void *stream, int dtype, int stochastic, const float *in, const uint32_t *rnd,
                  uint16_t *out, int64_t n);

/* FP64 reductions over the compact fields (cal_max_v, SP:399-402, plus the
 * flux and density sums used for permeability and convergence):
 *   out[0] sum rho        out[1..3] sum v       out[4] max |v|
 *   out[5] count of non-finite rho or v         out[6] sum vx over i_lo <= i <= i_hi
 *   out[7] number of cells in that slab range   out[8] sum rho over that slab range
 *   out[9] 0
 * Cells with a non-finite value count in out[5] only.  `out` is device
 * double[10]; workspace of lbm3d_reduce_workspace_bytes(). */
size_t lbm3d_reduce_workspace_bytes(int32_t n_cells);
int lbm3d_reduce(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho,
                 const float *v, int32_t i_lo, int32_t i_hi, double *out, void *workspace,
                 size_t workspace_bytes);

/* The same reductions over FP64 fields (the lbm3d_macros output of FP64
 * storage, LBM3D_F64), so that FP64 runs are not rounded to FP32 before they
 * are summed. */
int lbm3d_reduce64(void *stream, const lbm3d_params *p, const int32_t *cells, const double *rho,
                   const double *v, int32_t i_lo, int32_t i_hi, double *out, void *workspace,
                   size_t workspace_bytes);

/* lbm3d_reduce without a workspace argument: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight convergence loops, harnesses that forbid hipMalloc).  The
 * per-block partials live in a fixed 80 KiB device array inside the library
 * (one copy per GPU, held until the process ends); nothing is allocated.  The
 * partials are fully rewritten by each call, so any n_cells that lbm3d_reduce
 * accepts is accepted.  Same results as lbm3d_reduce, bit for bit.  Calls that
 * overlap on one GPU (concurrent streams) share the array and race: serialise
 * them. */
int lbm3d_reduce_noalloc(void *stream, const lbm3d_params *p, const int32_t *cells, const float *rho,
                         const float *v, int32_t i_lo, int32_t i_hi, double *out);


/* ---- batches: many samples of uneven size in one launch per step ----------
 *
 * No reference counterpart: the reference runs one grid per solver object
 * and its drivers loop over samples in Python (example_porous_medium.py).
 * Samples of a batch share the physical parameters and face modes of `p`
 * (s_v, s_o, force, bc_mode, rho_bc, u_bc); each has its own grid.  Their
 * fluid lists are laid end to end, sample s owning compact ids
 * [cell_begin, cell_begin + n_cells) with cell_begin a multiple of
 * LBM3D_BATCH_ALIGN, so the batch is one population array g[19][S] with
 * S = lbm3d_slot_stride(cells) and one cells / face_bits / rows array:
 *   cells[cell_begin + i]    dense C-order index inside sample s's own grid;
 *   rows[rows_begin + ...]   sample s's row table, built with
 *                            lbm3d_build_rows_at(..., id_offset = cell_begin);
 *   face_bits                lbm3d_build_face_bits of sample s written at
 *                            face_bits + cell_begin / 16;
 *   face_u[face_u_begin ...] sample s's pressure-face state, faces in face order.
 * In every batch call p->n_cells is the layout's total `cells` and p->nx,
 * ny, nz are not used.  Every sample's list is compacted with p->band_rows
 * (lbm3d_compact_fill with the sample's extents and the batch's band_rows),
 * and lbm3d_batch_desc records that order for the odd steps' row windows.
 * Padding ids between samples are never read or written.
 */
#define LBM3D_BATCH_ALIGN 512

typedef struct lbm3d_sample {
  int32_t nx, ny, nz;      /* the sample's grid */
  int32_t n_cells;         /* its fluid cells (lbm3d_compact_count) */
  int32_t cell_begin;      /* set by lbm3d_batch_layout */
  int32_t face_u_begin;    /* set by lbm3d_batch_layout */
  int64_t rows_begin;      /* set by lbm3d_batch_layout */
} lbm3d_sample;

typedef struct lbm3d_batch_sizes {
  int64_t cells;       /* compact ids spanned (a multiple of LBM3D_BATCH_ALIGN): p->n_cells of batch calls */
  int64_t row_words;   /* entries of the batch row table */
  int64_t face_u;      /* elements of face_u */
  int64_t desc_bytes;  /* bytes of the descriptor buffer */
} lbm3d_batch_sizes;

/* Launch shape of one parity of the batched step: `block` threads per
 * workgroup (128, 256 or 512); grid_per_cu 0 = one workgroup per tile of
 * `block` cells, otherwise a persistent grid of grid_per_cu workgroups per
 * compute unit walking the tiles; reverse 1 = tiles in descending order. */
typedef struct lbm3d_launch {
  int32_t block;
  int32_t grid_per_cu;
  int32_t reverse;
  int32_t reserved;
} lbm3d_launch;

/* Place n_samples samples (nx, ny, nz, n_cells filled in by the caller): sets
 * cell_begin, face_u_begin, rows_begin of each and the batch sizes.  Checks
 * each sample's extents against the face modes of p (host only). */
int lbm3d_batch_layout(const lbm3d_params *p, lbm3d_sample *samples, int32_t n_samples,
                       lbm3d_batch_sizes *sizes);

/* Fill the descriptor (per-sample geometry and the chunk -> sample map) in
 * host memory `host_desc` of sizes.desc_bytes; the caller copies it to a
 * device buffer `desc` that the batch calls read (host only). */
int lbm3d_batch_desc(const lbm3d_params *p, const lbm3d_sample *samples, int32_t n_samples,
                     void *host_desc, size_t bytes);

/* lbm3d_init over every sample of the batch; rho0 / u0 (SoA [3][cells]) in
 * batch ids, or NULL. */
int lbm3d_batch_init(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc,
                     int32_t n_samples, const int32_t *cells, void *face_u, const float *rho0,
                     const float *u0);

/* One time step of every active sample (active[s] != 0, or all when active
 * is NULL) in one launch; cfg[0] / cfg[1] the launch shapes of even / odd
 * steps, or NULL for the defaults.  A sample's results do not depend on cfg,
 * on the other samples or on the batch: they are bit-identical to lbm3d_step
 * on that sample alone. */
int lbm3d_batch_step(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc,
                     int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                     const uint32_t *face_bits, void *face_u, const uint8_t *active, int64_t t,
                     const lbm3d_launch *cfg);

/* `nsteps` consecutive lbm3d_batch_step calls from step t0. */
int lbm3d_batch_run(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc,
                    int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                    const uint32_t *face_bits, void *face_u, const uint8_t *active, int64_t t0,
                    int32_t nsteps, const lbm3d_launch *cfg);

/* lbm3d_batch_run over the cells [cell_begin, cell_end) of the batch only:
 * the samples that lie in that range advance by nsteps steps from t0, the
 * others are untouched.  cell_begin and cell_end are multiples of
 * LBM3D_BATCH_ALIGN (sample boundaries, or the batch's end).  Samples are
 * independent, so a host may advance groups of samples in turn, several
 * steps per visit, while each group's populations stay in the last-level
 * cache; every sample's results are bit-identical to lbm3d_step on it alone.
 * No reference counterpart (example_porous_medium.py:12-37 runs samples one
 * after another to convergence). */
int lbm3d_batch_run_range(void *stream, const lbm3d_params *p, int dtype, void *g, const void *desc,
                          int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                          const uint32_t *face_bits, void *face_u, const uint8_t *active, int64_t t0,
                          int32_t nsteps, int32_t cell_begin, int32_t cell_end, const lbm3d_launch *cfg);

/* lbm3d_macros over the batch: rho[cells], v SoA [3][cells] in batch ids. */
int lbm3d_batch_macros(void *stream, const lbm3d_params *p, int dtype, const void *g, const void *desc,
                       int32_t n_samples, const int32_t *cells, const lbm3d_row_word *rows,
                       const uint32_t *face_bits, const void *face_u, int64_t t, void *rho, void *v);

/* Per-sample FP64 reductions of FP32 batch macros: out (device double
 * [n_samples][8]) = sum rho, sum v (3), max |v|, non-finite cells, fluid
 * cells, 0. */
size_t lbm3d_batch_reduce_workspace_bytes(int64_t cells);
int lbm3d_batch_reduce(void *stream, const lbm3d_params *p, const void *desc, int32_t n_samples,
                       const float *rho, const float *v, double *out, void *workspace,
                       size_t workspace_bytes);

#ifdef __cplusplus
}
#endif

#endif /* LBM3D_H */
