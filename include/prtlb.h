/*
 * prtlb.h -- C ABI of libprtlb_hip.so, the MI355X (gfx950) PRT-LB hot path.
 *
 * The reference (shahram444/PRT-DeepONet-GUI) has no FFI: its hot path is two
 * NumPy functions, stokes_d2q9/stokes_d3q19 and solve_adr.  Each entry point
 * below replaces the NumPy loop it names; the Python package `prtlb`
 * (prt-deeponet-gui_amd/prtlb) binds them with ctypes and keeps the
 * reference's Python signatures.
 *
 * Conventions
 *   - every function returns 0 on success or a negative PRT_E* code; the
 *     message for the calling thread is prt_last_error().
 *   - all array arguments are DEVICE pointers owned by the caller (PyTorch
 *     tensors in the Python host); the library never allocates, frees or
 *     keeps them.  `stream` is a hipStream_t (NULL = the null stream); every
 *     launch is asynchronous on it.
 *   - grids are C-ordered (x slowest, last axis fastest), exactly the NumPy
 *     layout of the reference's (nx, ny[, nz]) arrays.  A 2D grid is passed
 *     with nz = 1.
 *   - dense grids store populations by z-row: the q populations of row
 *     (x, y) follow one another, each a contiguous run of zs elements along
 *     z, so population i of node (x, y, z) lives at
 *       f[((x * ny + y) * q + i) * zs + z],
 *     zs = nz rounded up to whole 128-byte lines (a 2D grid is the row set
 *     (nx, 1, ny)).  A buffer holds prt_lbm_pop_elems() elements.
 *   - fluid lists (prt_lbm_sparse_*) store populations by tile of 256
 *     consecutive list entries, the q populations of a tile one after the
 *     other: f[((k / 256) * q + i) * 256 + k % 256] for entry k.  A buffer
 *     holds q * prt_lbm_pop_stride(n_fluid, storage) elements.
 */
#ifndef PRTLB_H
#define PRTLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PRT_OK 0
#define PRT_E_ARG -1     /* bad argument (shape, null pointer, enum) */
#define PRT_E_HIP -2     /* a HIP runtime call failed */
#define PRT_E_RANGE -3   /* size beyond what the kernel indexes */

/* population storage (all collide arithmetic in FP64 / FP32 / FP32) */
#define PRT_STORE_FP64 0 /* f_i, double: the reference's own arithmetic */
#define PRT_STORE_FP32 1 /* f_i - w_i, float */
#define PRT_STORE_FP16 2 /* (f_i - w_i) * (s / w_i), IEEE half, stored with a
                            deterministic dither (rounded stochastically) */

/* prt_lbm_step / prt_lbm_run flags */
#define PRT_LBM_SKIP_SOLID 1 /* solid nodes are neither read nor written */
/* launch shape: 2^bz_log2 x 2^by_log2 threads (64..512), npt (1 or 2) nodes
 * per thread along y; 0 = the library's default shape */
#define PRT_LBM_TUNE(bz_log2, by_log2, npt) \
  (((bz_log2) << 8) | ((by_log2) << 12) | (((npt) - 1) << 16))
/* nontemporal population loads (dense mode, one node per thread) */
#define PRT_LBM_NT_LOADS (1 << 18)
/* XCD-contiguous block order (dense mode, one node per thread): each of the
 * 8 XCDs runs a contiguous slab of blocks, so lines shared by neighbouring
 * blocks meet in one L2; results are unchanged */
#define PRT_LBM_XCD (1 << 19)
/* prt_lbm_batch_step: the members are fluid lists (prt_lbm_segment.link) */
#define PRT_LBM_SPARSE (1 << 1)
/* prt_lbm_step / _run / _velocity on an x-slab (one GPU's share of a box
 * split along x): the grid's planes 0 and nx - 1 are halo planes, filled
 * by the caller before each step (prt_lbm_halo_unpack), and only planes
 * 1 .. nx - 2 are updated and reduced */
#define PRT_LBM_SLAB (1 << 20)

/* Message of the last failing call on this thread ("" if none). */
const char *prt_last_error(void);

/* Library version string. */
const char *prt_version(void);

/* sizeof(prt_adr_params) as compiled into the library (binding check). */
size_t prt_adr_params_size(void);

/* ------------------------------------------------------------------ Stokes */

/* The lattice the kernels stream and collide with, copied to host arrays:
 * c (3q int32, velocity i at c[3i .. 3i+2] in the reference's (x, y, z)
 * axes; D2Q9 has its y component in slot 1 and 0 in slot 2), w (q doubles)
 * and opp (q int32).  q = 9 follows 3D/tools/prtlb_2d.py:391-394, q = 19
 * 3D/tools/prtlb_3d.py:370-375, order included.  Host only. */
int prt_lbm_lattice(int q, int32_t *c, double *w, int32_t *opp);

/* Elements of one population buffer of a dense (nx, ny, nz) grid (z-row
 * layout above), or -1 for a bad grid. */
int64_t prt_lbm_pop_elems(int q, int storage, int nx, int ny, int nz);

/* Entries of an n_nodes fluid list rounded up to whole 256-entry tiles
 * (a population buffer holds q times this many elements). */
int64_t prt_lbm_pop_stride(int64_t n_nodes, int storage);

/* Rest state f_i = w_i (zero deviation in the shifted storages) over a
 * whole dense population buffer.  Replaces `f = np.ones(...) * w`
 * (3D/tools/prtlb_3d.py:403, 3D/tools/prtlb_2d.py:427). */
int prt_lbm_fill_rest(int q, int storage, int nx, int ny, int nz, void *f,
                      void *stream);

/* One fused TRT iteration: pull-stream from f_src, collide, Guo force,
 * solid-node reversal, write f_dst.  Replaces the loop body at
 * 3D/tools/prtlb_3d.py:427-445 (D3Q19, q = 19) and
 * 3D/tools/prtlb_2d.py:449-472 (D2Q9, q = 9).
 *   solid     n_nodes bytes, nonzero = solid (g != PORE)
 *   omega_p   1/tau; omega_m from the magic parameter (prt_trt_minus)
 *   force     body force along +x
 *   fp16_scale  s of PRT_STORE_FP16 (ignored otherwise)
 *   partials  if non-NULL: per-block FP64 partial sums of |u_x| over fluid
 *             nodes (the FlowConvergence quantity, 3D/tools/prtlb_3d.py:444)
 *             are written here; prt_lbm_partials_count() gives the length.
 *   f_dst must not alias f_src.  With PRT_LBM_SKIP_SOLID, f_dst must hold
 *   the state of the step before f_src (the A/B pair) -- bounce-back then
 *   reads a fluid node's own two-step-old slot instead of the solid node. */
int prt_lbm_step(int q, int storage, int nx, int ny, int nz,
                 const uint8_t *solid, const void *f_src, void *f_dst,
                 double omega_p, double omega_m, double force,
                 double fp16_scale, int flags, double *partials, void *stream);

/* Length a partials buffer needs for this grid (any launch shape). */
int64_t prt_lbm_partials_count(int q, int nx, int ny, int nz);

/* Partials one launch with these flags writes (indices [0, blocks)); the
 * velocity kernel launches the flags = 0 shape.  Sum exactly these. */
int64_t prt_lbm_launch_blocks(int q, int nx, int ny, int nz, int flags);

/* `steps` fused iterations alternating f_a -> f_b -> f_a ...; the last one
 * writes f_b if steps is odd, else f_a.  Partials (if non-NULL) are
 * produced by the last iteration only. */
int prt_lbm_run(int q, int storage, int nx, int ny, int nz,
                const uint8_t *solid, void *f_a, void *f_b, int steps,
                double omega_p, double omega_m, double force,
                double fp16_scale, int flags, double *partials, void *stream);

/* Deterministic FP64 sum of n doubles (fixed order) into *out (device). */
int prt_sum_f64(const double *x, int64_t n, double *out, void *stream);

/* prt_lbm_run (steps >= 1) followed by prt_sum_f64 of the last iteration's
 * per-block |u_x| partials into *sum_out (device): the FlowConvergence tick,
 * with the same bits as those two calls (partials summed over
 * prt_lbm_launch_blocks(q, nx, ny, nz, flags) entries), and no allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops).  The partials live in the
 * library's own device array of prt_lbm_static_partials_max() doubles (one
 * per GPU, held until the process ends); a launch needing more returns
 * PRT_E_RANGE.  Calls that overlap on one GPU share that array and race. */
int64_t prt_lbm_static_partials_max(void);
int prt_lbm_run_sum_static(int q, int storage, int nx, int ny, int nz,
                           const uint8_t *solid, void *f_a, void *f_b,
                           int steps, double omega_p, double omega_m,
                           double force, double fp16_scale, int flags,
                           double *sum_out, void *stream);

/* Velocity from the post-collision state f_last (pull, density, Guo-shifted
 * momentum, zero in solid).  Replaces 3D/tools/prtlb_3d.py:451-455 and
 * 3D/tools/prtlb_2d.py:481-485.  With PRT_LBM_SKIP_SOLID in flags, f_prev is
 * the other buffer of the pair (the state one iteration older), as
 * prt_lbm_step's fluid-only mode requires; otherwise it may be NULL.
 * v64 (3, n) double and/or v32 (3, n) float, either may be NULL; for 2D the
 * third row is 0.  partials (device, may be NULL, prt_lbm_partials_count()
 * long) receives per-block FP64 sums of u_x over every node, for the Darcy
 * permeability (reduce with prt_sum_f64). */
int prt_lbm_velocity(int q, int storage, int nx, int ny, int nz,
                     const uint8_t *solid, const void *f_last,
                     const void *f_prev, int flags, double force,
                     double fp16_scale, double *v64, float *v32,
                     double *partials, void *stream);

/* Sum of u_x over each x-plane of v64 (3, nx, ny, nz): plane[nx]. */
int prt_plane_flux(const double *v64, int nx, int ny, int nz, double *plane,
                   void *stream);

/* ------------------------------------------------ fluid-list (sparse) LBM */
/* The same TRT step over the fluid nodes only, a sparse pore-node list (the
 * reference steps every box node, 3D/tools/prtlb_3d.py:403-445).  fluid_idx lists the n_fluid fluid (non-solid) box nodes in C
 * order; populations are tiled by 256 entries (see the conventions at the
 * top; q * prt_lbm_pop_stride(n_fluid, storage) elements); link[(i-1) * lk + k],
 * lk = prt_lbm_link_stride(n_fluid), is the entry that streams into entry k
 * along e_i (periodic) or -1 where that node is solid.  Results equal the
 * dense PRT_LBM_SKIP_SOLID ones bit for bit. */
int64_t prt_lbm_link_stride(int64_t n_fluid);
/* Build the link table; map_work: n_box int32 of scratch. */
int prt_lbm_link_table(int q, int nx, int ny, int nz, const int32_t *fluid_idx,
                       int64_t n_fluid, int32_t *map_work, int32_t *link,
                       void *stream);
int prt_lbm_sparse_fill_rest(int q, int storage, int64_t n_fluid, void *f,
                             void *stream);
/* Partials length of the sparse run / velocity (one per 256 entries). */
int prt_lbm_sparse_partials_count(int64_t n_fluid);
/* As prt_lbm_run over the list (f_b must hold the state before f_a). */
/* The link table as 16-bit steps: entry s of link[(i-1) * lk + k] becomes
 * s - k taken modulo n_fluid into [-32767, 32767], and -32768 marks a solid
 * upstream node.  Halves the table's bytes (and its share of every step's
 * traffic: FP32 224 -> 188 B per fluid node).  link16 holds (q-1) * lk
 * int16; *overflow (a device int32) becomes 1 if some step does not fit
 * (a list with more than ~32k fluid nodes per x-plane): then keep using the
 * 32-bit table. */
int prt_lbm_link_pack16(int q, int64_t n_fluid, const int32_t *link,
                        int16_t *link16, int32_t *overflow, void *stream);
/* link_bits: 32 for the prt_lbm_link_table form, 16 for the packed one. */
int prt_lbm_sparse_run(int q, int storage, int64_t n_fluid, const void *link,
                       int link_bits, void *f_a, void *f_b, int steps,
                       double omega_p, double omega_m, double force,
                       double fp16_scale, double *partials, void *stream);
/* As prt_lbm_velocity (fluid-only) over the list: writes the fluid entries
 * of v64 / v32 (3, n_box) at their box nodes; the caller zeroes the rest.
 * partials[b]: FP64 sum of u_x per block. */
int prt_lbm_sparse_velocity(int q, int storage, int64_t n_box, int64_t n_fluid,
                            const int32_t *fluid_idx, const void *link,
                            int link_bits, const void *f_last,
                            const void *f_prev, double force,
                            double fp16_scale, double *v64, float *v32,
                            double *partials, void *stream);

/* --------------------------------------------------------------- transport */

/* Scalar parameters of one transport solve, built by the host exactly as
 * 3D/tools/prtlb_3d.py:831-1018 builds them.  Every value the float32 path
 * uses is a float32 number carried in a double (lossless). */
typedef struct prt_adr_params {
  int32_t nd;            /* 2 or 3 */
  int32_t ns;            /* chemicals, 1..8 */
  int32_t biotic_on;     /* 3D/tools/prtlb_3d.py:1288 */
  int32_t coupled;       /* biomass_coupled && ns > 3 && c_init[3] > 0 */
  int32_t abio_index;    /* abiotic reactant, or -1 when the abiotic rate is
                            the scalar zero (ka <= 0 or ns <= index) */
  int32_t abio_target;   /* abiotic_index if < ns, else -1: where -Ra lands */
  int32_t abio_product;  /* product index (!= reactant, < ns) or -1 */
  int32_t surface_only;  /* abiotic_surface_only */
  int32_t sorption;      /* 1: reversible sorption (build-defined) between
                            mobile sorb_c and immobile sorb_s */
  int32_t sorb_c, sorb_s;
  int32_t write_solid;   /* 1: also write the zeros outside the pore */
  double dt, k, ka, kd;
  double ks_donor, ks_acceptor;  /* <= 0 means first order */
  double alpha_dt;       /* float(alpha) * dt, as NumPy rounds it */
  double neg_alpha;      /* -alpha */
  double alpha_pos;      /* 1 if alpha > 0 */
  double c2;             /* 1 - yield (ns > 3) or 1, rounded once */
  double yield_;
  double c3_init;
  double k_sorb, k_desorb; /* reversible sorption rates */
  int32_t mobile[8];     /* 1 = transported */
  int32_t pinned[8];     /* 1 = mobile with a feed: held at fed[j] on x = 0 */
  double fed[8];
  double dface[8][4];    /* harmonic face diffusivity, [j][2*bf_self + bf_nb] */
  int32_t immobile_list; /* pore-list launches: the rows of immobile
                            chemicals hold the pore voxels in list order
                            (entry k at k) instead of box order; their
                            loads and stores are then contiguous.  The
                            dense launches refuse it. */
  int32_t reserved;
} prt_adr_params;

/* One explicit step of the flux-form MUSCL-minmod transport plus reactions:
 * inlet pin (x = 0, fed mobile species), transport of the mobile rows,
 * C + dt (T + R), clip at zero in the pore, zero outside.  Replaces step()
 * at 3D/tools/prtlb_3d.py:1392-1413 with transport :1182-1208 and reaction
 * :1288-1367 fused in.  Grid (nx, ny, nz) with nz = 1 for 2D.
 *   flags  n bytes: bit0 pore-or-biofilm (solute may be here), bit1 biofilm
 *   u      (nd, n): velocity normalised by its axial pore mean, 0 off the
 *          pore (3D/tools/prtlb_3d.py:860-865)
 *   c_in, c_out  (ns, n); c_out must not alias c_in.  Unless write_solid,
 *          voxels outside the pore are not written (keep them zero).
 * The float32 path is bit-exact with the reference's float32 NumPy. */
int prt_adr_step(const prt_adr_params *p, int nx, int ny, int nz,
                 const uint8_t *flags, const float *u, const float *c_in,
                 float *c_out, void *stream);

/* Pore table of a pore list, built once per transport solve (it depends on
 * the flags and the velocity only): one device buffer of
 * prt_adr_table_bytes(nd, fp64, n_pore) bytes (16-byte aligned).  pore_idx
 * holds n_pore flat voxel indices in C order, the order the list steps
 * need.  For entry k the table records the voxel index, its word (the reach
 * codes and face bits of that voxel) and its face velocities, the
 * reference's open-face averages (3D/tools/prtlb_3d.py:1100-1135) computed
 * with the same operations as prt_adr_step, in aligned per-entry records
 * (the layout is the library's own).  An entry that is not a pore voxel, or
 * not below nx*ny*nz, is marked and skipped by the list steps.  Returns -1
 * (and an error message) for a bad nd or count. */
int64_t prt_adr_table_bytes(int nd, int fp64, int64_t n_pore);
int prt_adr_pore_table(int nd, int nx, int ny, int nz, const uint8_t *flags,
                       const float *u, const int32_t *pore_idx, int64_t n_pore,
                       void *pore_table, void *stream);
int prt_adr_pore_table_f64(int nd, int nx, int ny, int nz,
                           const uint8_t *flags, const double *u,
                           const int32_t *pore_idx, int64_t n_pore,
                           void *pore_table, void *stream);

/* `steps` float32 steps with one lane per pore-list entry, geometry from the
 * pore table of the same list and grid: no lane works on solid,
 * and no flag or velocity grid is read.  Same arithmetic, same bits as
 * prt_adr_run on the listed voxels; voxels outside the list are never
 * written, so write_solid must be 0 (PRT_E_ARG). */
int prt_adr_run_list(const prt_adr_params *p, int nx, int ny, int nz,
                     const void *pore_table, int64_t n_pore, float *c_a,
                     float *c_b, int steps, void *stream);

/* One FP64 step over the pore list (FP64 pore table). */
int prt_adr_step_list_f64(const prt_adr_params *p, int nx, int ny, int nz,
                          const void *pore_table, int64_t n_pore,
                          const double *c_in, double *c_out, void *stream);

/* Same step in FP64 (the build's own double-precision variant). */
int prt_adr_step_f64(const prt_adr_params *p, int nx, int ny, int nz,
                     const uint8_t *flags, const double *u,
                     const double *c_in, double *c_out, void *stream);

/* `steps` float32 steps alternating c_a -> c_b -> c_a ...; the last one
 * writes c_b if steps is odd, else c_a. */
int prt_adr_run(const prt_adr_params *p, int nx, int ny, int nz,
                const uint8_t *flags, const float *u, float *c_a, float *c_b,
                int steps, void *stream);

/* ------------------------------------------------------------- x-slabs */

/* Halo exchange of a box split along x into slabs (the layout one GPU of a
 * multi-GPU solve holds; no counterpart in the reference, which solves one
 * box in one process).  A slab of m planes is stepped as the grid
 * (m + 2, ny, nz) with PRT_LBM_SLAB; before each step its plane 0 must hold
 * the populations with e_x = +1 of the left neighbour's last own plane and
 * its plane m + 1 those with e_x = -1 of the right neighbour's first own
 * plane, from the neighbour's current source buffer.  Each message is
 * prt_lbm_halo_elems() elements: the populations with e_x = dir (5 for
 * D3Q19, 3 for D2Q9, in lattice order), each as the plane's ny row runs of
 * zs elements (the z-row layout's own runs, padding included).
 * prt_lbm_halo_pack copies them out of plane `plane` of f, unpack writes a
 * message into plane `plane`.  (nx, ny, nz) is the slab's grid with its
 * halo planes; dir is +1 or -1.  The message moves between GPUs with
 * whatever the caller uses (RCCL send/recv, a peer copy). */
int64_t prt_lbm_halo_elems(int q, int storage, int nx, int ny, int nz);
int prt_lbm_halo_pack(int q, int storage, int nx, int ny, int nz,
                      const void *f, int plane, int dir, void *buf,
                      void *stream);
int prt_lbm_halo_unpack(int q, int storage, int nx, int ny, int nz,
                        void *f, int plane, int dir, const void *buf,
                        void *stream);

/* ------------------------------------------------------------- batching */

/* Many geometries of uneven size advanced by ONE launch (the batched
 * campaign and the serving loop).  No counterpart in the reference, which
 * solves one geometry per process; each segment is one call of
 * stokes_d2q9/stokes_d3q19 (or one solve_adr step) and gives bit-identical
 * results alone or in any batch. */
typedef struct prt_lbm_segment {
  const uint8_t *solid;   /* n bytes, device */
  void *f_a, *f_b;        /* the segment's two population buffers (prt_lbm_pop_elems) */
  int32_t nx, ny, nz;     /* nz = 1 for D2Q9 */
  int32_t phase;          /* which buffer is newest at launch parity 0:
                             0 = f_a, 1 = f_b */
  double omega_p, omega_m, force, fp16_scale;
  const void *link;       /* fluid-list member (PRT_LBM_SPARSE batches): its
                             link table (prt_lbm_link_table, or its 16-bit
                             form from prt_lbm_link_pack16); f_a / f_b then
                             hold n_fluid entries per population; NULL for a
                             dense member.  All members of a batch agree. */
  int64_t n_fluid;
  int32_t link_bits;      /* 32 (or 0) or 16: width of the link table */
  int32_t reserved;
} prt_lbm_segment;

typedef struct prt_adr_segment {
  const prt_adr_params *params;   /* host pointer, read by the plan call */
  int32_t nx, ny, nz, phase;
  const void *pore_table;         /* device pore table (prt_adr_pore_table) */
  int64_t n_pore;                 /* entries of its C-ordered pore list */
  float *c_a, *c_b;               /* device concentration buffers */
} prt_adr_segment;

/* Bytes of one device descriptor of each kind. */
int64_t prt_lbm_batch_desc_size(void);
int64_t prt_adr_batch_desc_size(void);

/* Build the descriptors (desc_host: nseg * desc_size bytes, HOST) and the
 * block-to-segment table (blk_seg_host, HOST, NULL = size query into
 * *n_blocks); the caller copies both to the device. */
int prt_lbm_batch_plan(int q, int storage, const prt_lbm_segment *segs, int nseg,
                       void *desc_host, int32_t *blk_seg_host,
                       int64_t blk_capacity, int64_t *n_blocks);
int prt_adr_batch_plan(const prt_adr_segment *segs, int nseg, void *desc_host,
                       int32_t *blk_seg_host, int64_t blk_capacity,
                       int64_t *n_blocks, int *max_species);

/* One iteration of every segment.  parity flips each launch; partials
 * (device, n_blocks doubles; NULL = none) receive per-block sums of |u_x|
 * over fluid nodes. */
int prt_lbm_batch_step(int q, int storage, const void *desc_dev,
                       const int32_t *blk_seg_dev, int64_t n_blocks,
                       int parity, int flags, double *partials, void *stream);

/* One transport step of every segment; partials (NULL = none) receive per
 * block, per chemical slot, FP64 pore sums of the new field: width is 1, 2,
 * 4 or 8 (the smallest holding max_species). */
int prt_adr_batch_step(const void *desc_dev, const int32_t *blk_seg_dev,
                       int64_t n_blocks, int max_species, int parity,
                       double *partials, void *stream);

/* out[s * width + j] = sum of partials[b * width + j] over the blocks
 * [blk0[s], blk0[s] + nblk[s]) of segment s, in a fixed order. */
int prt_seg_sums(const int32_t *blk0_dev, const int32_t *nblk_dev, int nseg,
                 int width, const double *partials, double *out,
                 void *stream);

/* Launch shapes of the batched steps, the choices of the library's
 * launch-shape tuner (prtlb.autotune): 2^t_log2 threads per block (64, 128
 * or 256), 2^e_log2 runs of list entries per thread (1, 2 or 4; thread t of
 * a block takes entries base + e * threads + t), t_log2 + e_log2 = 8 or 9
 * (a block covers a tile of 256 or 512 entries), and the block order
 * (xcd = 1: each of the 8 XCDs runs one contiguous range of blocks; 0:
 * round robin).  Dense (box) Stokes members take only 256 x 1, either
 * order; transport partial sums only 256 x 1.  A batch is planned and
 * stepped with the same shape: a step of another shape than its plan's
 * updates nothing.  Every shape gives the same bits (the same per-entry
 * code; partial sums per 256-entry tile in one fixed order).  No
 * counterpart in the reference. */
#define PRT_BATCH_SHAPE(t_log2, e_log2, xcd) \
  ((t_log2) | ((e_log2) << 4) | ((xcd) << 8))
/* the shapes of prt_lbm_batch_plan/_step and prt_adr_batch_plan/_step */
#define PRT_BATCH_SHAPE_LBM_DEFAULT PRT_BATCH_SHAPE(8, 0, 1)
#define PRT_BATCH_SHAPE_ADR_DEFAULT PRT_BATCH_SHAPE(8, 0, 0)

/* Entries per block of a shape, or 0 for a code that is not a shape. */
int prt_batch_shape_tile(int shape);

/* prt_lbm_batch_plan / prt_adr_batch_plan for a given shape.  *n_tiles
 * (may be NULL) receives the partial-sum slots a step writes: one per 256
 * entries of each member, sum over members of ceil(n / 256) (times the
 * width for transport); blk0 / nblk of prt_seg_sums count these tiles.
 * At a 256-entry shape n_tiles equals n_blocks. */
int prt_lbm_batch_plan_shaped(int q, int storage, const prt_lbm_segment *segs,
                              int nseg, int shape, void *desc_host,
                              int32_t *blk_seg_host, int64_t blk_capacity,
                              int64_t *n_blocks, int64_t *n_tiles);
int prt_adr_batch_plan_shaped(const prt_adr_segment *segs, int nseg, int shape,
                              void *desc_host, int32_t *blk_seg_host,
                              int64_t blk_capacity, int64_t *n_blocks,
                              int64_t *n_tiles, int *max_species);

/* prt_lbm_batch_step / prt_adr_batch_step with the plan's shape; partials
 * (NULL = none) then hold n_tiles slots (times the width). */
int prt_lbm_batch_step_shaped(int q, int storage, const void *desc_dev,
                              const int32_t *blk_seg_dev, int64_t n_blocks,
                              int parity, int flags, int shape,
                              double *partials, void *stream);
int prt_adr_batch_step_shaped(const void *desc_dev, const int32_t *blk_seg_dev,
                              int64_t n_blocks, int max_species, int parity,
                              int shape, double *partials, void *stream);

/* ------------------------------------------------------------- geometry */

/* Seeded grain packs on the GPU.  The host draws the candidate grains
 * (centres (n_grains, nd) and radii, in generator order) and bins them into
 * cubic cells of edge cell_edge (CSR: cell_start[ncells + 1], cell_items,
 * ascending grain order per cell).  first[n] = index of the first candidate
 * whose periodic minimum-image distance to voxel n is <= its radius, or
 * n_grains.  Grid (nx, ny, nz), nz = 1 for 2D.  Build-defined workload
 * generator (the reference's generators threshold Gaussian fields). */
int prt_grain_first_cover(int nx, int ny, int nz, const double *centres,
                          const double *radii, int n_grains,
                          const int32_t *cell_start, const int32_t *cell_items,
                          int cell_edge, int32_t *first, void *stream);

/* hist[k] += number of voxels whose first cover is k (hist zeroed by the
 * caller, n_grains + 1 entries). */
int prt_grain_histogram(const int32_t *first, int64_t n_voxels, int32_t *hist,
                        void *stream);

/* g[n] = first[n] < n_grains_used ? solid_code : pore_code. */
int prt_grain_mask(const int32_t *first, int64_t n_voxels, int n_grains_used,
                   int solid_code, int pore_code, uint8_t *g, void *stream);

/* Pore voxels not face-connected (no wrap) to a pore voxel on x = 0 become
 * solid_code, in place; *removed_host = their count.  reach: n bytes,
 * scratch2: 2 ints (device).  Synchronises the stream.  Replaces
 * keep_inlet_connected, 3D/tools/prtlb_3d.py:530-564. */
int prt_keep_inlet_connected(uint8_t *g, int nx, int ny, int nz, int pore_code,
                             int solid_code, uint8_t *reach, int *scratch2,
                             int max_sweeps, int *removed_host, void *stream);

/* Pore voxels within `layers` face dilations (no wrap) of the non-pore get
 * `code`, in place; *marked_host = their count.  work_a, work_b: n bytes;
 * scratch1: 1 int.  Synchronises the stream.  Replaces coat_with_biofilm,
 * 3D/tools/prtlb_3d.py:567-593. */
int prt_coat_with_biofilm(uint8_t *g, int nx, int ny, int nz, int pore_code,
                          int code, int layers, uint8_t *work_a,
                          uint8_t *work_b, int *scratch1, int *marked_host,
                          void *stream);

/* Geodesic distance from the inlet through the pore, the relaxation of
 * 3D/tools/build_practice_dataset.py:53-72 (geodesic) pass for pass: float32,
 * solid -> 0 in out, unreached pore keeps 1e9.  d_a, d_b: n floats of work;
 * scratch1: one int.  Runs at most max_sweeps sweeps (the reference uses
 * 4 nx) and reports the sweeps run (a multiple of 8 or max_sweeps).
 * Synchronises the stream. */
int prt_geodesic(const uint8_t *g, int nx, int ny, int nz, int pore_code,
                 float *d_a, float *d_b, int *scratch1, int max_sweeps,
                 float *out, int *sweeps_host, void *stream);

/* Euclidean distance of every pore voxel to the nearest non-pore voxel,
 * float32, 0 on non-pore: scipy.ndimage.distance_transform_edt(g == PORE)
 * as 3D/tools/build_dataset_3d.py:659 stores it.  work_a, work_b: n int32.
 * The grid must hold at least one non-pore voxel (the caller checks); edges
 * above 16384 are refused. */
int prt_edt(const uint8_t *g, int nx, int ny, int nz, int pore_code,
            int32_t *work_a, int32_t *work_b, float *out, void *stream);

/* --------------------------------------------- harmonic pressure prior */
/* Laplace's equation on the pore (byte mask, 1 = pore) face-connected to
 * the inlet: P = 1 on x = 0, 0 on x = nx - 1, no flux through walls; 0 on
 * solid and on pore the inlet does not reach.  Replaces harmonic_pressure
 * (3D/tools/harmonic_pressure.py:91-146; its scipy spsolve becomes
 * Jacobi-preconditioned CG in FP64 to relative residual tol, then float32).
 * A 2D grid is passed as (nx, 1, ny).  work: prt_harmonic_work_size()
 * bytes of device scratch.  *dropped_host (if non-NULL): the unreached pore
 * voxels; *iters_host: CG iterations.  Synchronises the stream. */
int64_t prt_harmonic_work_size(int nx, int ny, int nz);
int prt_harmonic_pressure(const uint8_t *pore, int nx, int ny, int nz,
                          void *work, int64_t work_bytes, double tol,
                          int max_iter, float *out, int *dropped_host,
                          int *iters_host, void *stream);
/* Its gradient (harmonic_gradient, :148-197): out (nd, n) float32, central
 * differences where both neighbours are pore, one-sided where one is, 0
 * where none is; with normalise, divided by the largest |component| over
 * the pore.  scratch: ceil(n / 256) + 1 floats. */
int prt_harmonic_gradient(const uint8_t *pore, int nx, int ny, int nz, int nd,
                          const float *P, int normalise, float *out,
                          float *scratch, void *stream);

/* --------------------------------------------- flow descriptors */
/* Pore descriptors of 3D/tools/flow_features.py, bit-identical float32
 * maps; pore is a byte mask (1 = pore), edt its prt_edt distances
 * (pore_code 1); a 2D grid is (nx, 1, ny).
 * UPRM (:205-257): the widest path from the inlet; best: n floats;
 * scratch1: one device int.  Synchronises the stream. */
int prt_uprm(const uint8_t *pore, int nx, int ny, int nz, const float *edt,
             float *best, int *scratch1, int max_sweeps, int *sweeps_host,
             void *stream);
/* dw2 (:260-280): clip((edt^2 - lo) / span, 0, 1) on the pore, float32. */
int prt_dw2(const uint8_t *pore, int64_t n, const float *edt, float lo,
            float span, float *out, void *stream);
/* Local thickness (:128-150) with the sweep numpy.arange(rmax, 0, -r_step),
 * rmax the largest edt; rc_work: n doubles. */
int prt_local_thickness(const uint8_t *pore, int nx, int ny, int nz,
                        const float *edt, double rmax, double r_step,
                        double *rc_work, float *out, void *stream);
/* Painting into the buffers (:153-177) and the MIS finish (:180-201): full
 * holds the interior slab's local thickness (x in [buf, nx - buf)) and is
 * completed in place; vmax, vmin: the largest and smallest positive
 * interior values. */
int prt_paint_buffers(const uint8_t *pore, int nx, int ny, int nz, int buf,
                      float vmax, float vmin, float *full, void *stream);

/* ------------------------------------------------ check_physics passes */
/* The data passes of check_physics (3D/tools/prtlb_3d.py:1632-1843) over
 * device snapshots out (n_t, ns, n_vox) float32 with a pore byte mask
 * (1 = pore).  Each writes one partial per block (`blocks` from
 * prt_check_blocks); the caller folds the partials. */
int prt_check_blocks(int64_t n);
/* counts[2b], counts[2b+1]: NaNs and infinities of x[0 .. n) */
int prt_check_nonfinite(const float *x, int64_t n, int blocks, int64_t *counts,
                        void *stream);
/* mx/mn[(t*ns + j)*blocks + b]: pore max / min of out[t, j] (-inf / +inf
 * for a block without pore) */
int prt_check_pore_range(const float *out, int n_t, int ns, int64_t n_vox,
                         const uint8_t *pore, int blocks, float *mx, float *mn,
                         void *stream);
/* part[(p*n_t + t)*blocks + b]: max of chemical 0 of snapshot t over the x
 * plane p (0: x = nx/2, 1: x = nx-1), plane = voxels per x plane */
int prt_check_plane_max(const float *out, int n_t, int ns, int nx, int64_t plane,
                        int blocks, float *part, void *stream);
/* counts[2b], counts[2b+1]: pore triples along x and sign flips among them
 * in the field f (the last snapshot of chemical 0) */
int prt_check_checker(const float *f, const uint8_t *pore, int nx, int64_t plane,
                      int blocks, int64_t *counts, void *stream);
/* mom[((j-1)*blocks + b)*5 + q], j = 1 .. ns-1: FP64 sums over (snapshot,
 * pore voxel) of a0, aj, (a0-m0)^2, (aj-mj)^2, (a0-m0)(aj-mj), m = mean
 * (device, ns doubles) or 0 when mean is NULL */
int prt_check_moments(const float *out, int n_t, int ns, const uint8_t *pore,
                      int64_t n_vox, int blocks, const double *mean,
                      double *mom, void *stream);
/* part[b]: max |aj - sc a0| in float32 over (snapshot, pore voxel) */
int prt_check_copy_residual(const float *out, int n_t, int ns, int j, float sc,
                            const uint8_t *pore, int64_t n_vox, int blocks,
                            float *part, void *stream);

/* ---------------------------------------------------------- solver handles */
/* Handles over the stateless calls above, for hosts that prefer an object
 * (the create / init / step_n / reduce / velocity / destroy set a non-Python
 * caller of the Stokes and transport stages binds).  A handle keeps the
 * caller's device pointers, grid, rates, stream and which buffer is newest;
 * it owns no device memory, is bound to one stream and is not shared between
 * threads.  Same arithmetic, same bits as the calls it wraps. */
typedef struct prt_lbm_solver prt_lbm_solver;
typedef struct prt_adr_solver prt_adr_solver;

#define PRT_REDUCE_MEAN_ABS_UX 0 /* mean |u_x| over fluid nodes after the last
                                    step_n (the FlowConvergence quantity) */
#define PRT_REDUCE_SUM_UX 1      /* sum of u_x over every node (Darcy) */

/* omega- of the TRT model: 1 / (magic / (1/omega+ - 1/2) + 1/2),
 * 3D/tools/prtlb_3d.py:130-139; NaN (and an error message) for tau <= 1/2. */
double prt_trt_minus(double omega_p, double magic);

/* partials: prt_lbm_partials_count() doubles; scalar: 1 double (device);
 * n_fluid: the number of fluid nodes (for the mean). */
int prt_lbm_solver_create(int q, int storage, int nx, int ny, int nz,
                          const uint8_t *solid, long long n_fluid, void *f_a,
                          void *f_b, double *partials, double *scalar,
                          double tau, double magic, double force,
                          double fp16_scale, int flags, void *stream,
                          prt_lbm_solver **out);
int prt_lbm_solver_init(prt_lbm_solver *h);          /* rest state, both buffers */
int prt_lbm_solver_step_n(prt_lbm_solver *h, int n); /* partials on the last */
/* Synchronises the stream; the value lands in *out_host. */
int prt_lbm_solver_reduce(prt_lbm_solver *h, int what, double *out_host);
int prt_lbm_solver_velocity(prt_lbm_solver *h, double *v64, float *v32);
void *prt_lbm_solver_newest(prt_lbm_solver *h);     /* newest population buffer */
int prt_lbm_solver_destroy(prt_lbm_solver *h);

/* A transport solve over the pore table of a pore list
 * (prt_adr_pore_table); the parameter block is copied into the handle. */
int prt_adr_solver_create(const prt_adr_params *p, int nx, int ny, int nz,
                          const void *pore_table, long long n_pore, float *c_a,
                          float *c_b, void *stream, prt_adr_solver **out);
int prt_adr_solver_step_n(prt_adr_solver *h, int n);
float *prt_adr_solver_newest(prt_adr_solver *h);    /* newest (ns, n) field */
/* Device-to-device copy of the newest field into dst (ns * n floats). */
int prt_adr_solver_snapshot(prt_adr_solver *h, float *dst);
int prt_adr_solver_destroy(prt_adr_solver *h);

/* --------------------------------------------------------- measurement */

/* Plain device-to-device copy of `bytes` (multiple of 16, 16-byte aligned
 * buffers): the STREAM-copy bandwidth ceiling and the known-size kernel for
 * calibrating HBM traffic counters.  variant 0: 16 B per lane, full grid;
 * 1: 64 B per lane; 2: grid-stride with nontemporal stores.  No counterpart
 * in the reference. */
int prt_stream_copy(const void *src, void *dst, int64_t bytes, int variant,
                    void *stream);

/* ------------------------------------------------------ NumPy-order sums */

/* Host-side plan of NumPy's float32 np.add.reduce order for n contiguous
 * values: 8192-element buffers summed left to right, each buffer by pairwise
 * recursion (n/2 rounded down to a multiple of 8 per split, leaves of at
 * most 128 with eight accumulators).  With plan == NULL only the sizes are
 * returned; the plan needs 2*n_leaves + 2*n_nodes + n_levels + 1 +
 * ceil(n/8192) int32 (HOST memory).  Used for the pore means of
 * 3D/tools/prtlb_3d.py:1500 (C[j][m].mean()). */
int prt_pairwise_plan(int64_t n, int32_t *plan, int64_t capacity,
                      int64_t *n_leaves, int64_t *n_nodes, int64_t *n_levels);

/* out[r] = the float32 NumPy sum of c[r * ld + idx[k]], k < n, for each of
 * `rows` rows, evaluated on the device with the plan (copied to the device)
 * and scratch of rows * (n_leaves + n_nodes) floats. */
int prt_pairwise_sum_f32(const float *c, int64_t ld, int rows,
                         const int32_t *idx, int64_t n, const int32_t *plan,
                         int64_t n_leaves, int64_t n_nodes, int64_t n_levels,
                         float *scratch, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PRTLB_H */
