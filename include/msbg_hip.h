/*
 * msbg_hip.h -- C ABI of the MI355X (gfx950) sparse-block-grid smoother and
 * particle-transfer path.
 *
 * Every entry point returns an int status (0 = ok, < 0 = error; the message
 * is available from msbg_last_error()).  No entry point unwinds or aborts;
 * the preconditions the reference only debug-asserts are checked here.
 * Pointers documented as "host" are read or written synchronously and not
 * retained; pointers documented as "device" must be device memory (from
 * msbg_malloc, hipMalloc or a torch tensor on the same device).  All work runs
 * on the library stream (msbg_set_stream, default the null stream); calls are
 * synchronous unless their name ends in _async.
 *
 * Reference interfaces are cited as <file>:<line> of msbg-rs.
 *
 * Layout of a grid level with block edge B = 2^L (L = 4 for 16^3 blocks):
 *   nx = ceil(sx / B) (sparse_grid.rs:131-136), bid = bx + by*nx + bz*nx*ny,
 *   vid = vx | vy << L | vz << 2L (sparse_grid.rs:159-175);
 *   b2s[bid] : 0 = empty dummy (absent block), 1 = full dummy, >= 2 = slot of
 *              an allocated value block;
 *   slab[slot * B^3 + vid] : payload; slots 0 / 1 hold the empty / full value.
 */
#ifndef MSBG_HIP_H
#define MSBG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSBG_OK = 0,
    MSBG_EINVAL = -1, /* invalid argument / violated precondition */
    MSBG_ENOMEM = -2, /* block capacity or device memory exhausted */
    MSBG_EHIP = -3,   /* HIP runtime error */
};

/* Stencil -- solver.rs:43-51. */
enum { MSBG_LAPLACIAN = 0, MSBG_MEAN_CURVATURE = 1, MSBG_BILAPLACIAN = 2 };
/* BoundaryCondition -- math/boundary.rs:6-14. */
enum { MSBG_BC_CLAMP = 0, MSBG_BC_NEUMANN = 1, MSBG_BC_DIRICHLET = 2 };

/* ---- runtime -------------------------------------------------------------- */
const char *msbg_last_error(void);
int msbg_device_count(int *n);
int msbg_set_stream(void *hip_stream); /* NULL = default stream */
int msbg_synchronize(void);
int msbg_malloc(void **dev_ptr, size_t bytes);
int msbg_free(void *dev_ptr);
int msbg_memcpy_htod(void *dev_dst, const void *host_src, size_t bytes);
int msbg_memcpy_dtoh(void *host_dst, const void *dev_src, size_t bytes);
int msbg_memcpy_dtod(void *dev_dst, const void *dev_src, size_t bytes);
/* Fills n_words 32-bit words of device memory with `value` (a kernel on the
 * library stream, asynchronous; e.g. a zero field without a host upload).  No
 * reference counterpart: device-memory utility of the C ABI. */
int msbg_fill_u32(void *dev_ptr, int64_t n_words, uint32_t value);
/* Timing with HIP events on the library stream. */
int msbg_event_create(void **ev);
int msbg_event_destroy(void *ev);
int msbg_event_record(void *ev);
int msbg_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- sparse grid -----------------------------------------------------------
 * A device-resident SparseGrid<f32, B, B^3> (sparse_grid.rs:61-155) or one
 * multires LevelData level (multires/level.rs:84-243).  The block pool is a
 * fixed-capacity slab (blockpool.rs:73-195): allocations beyond
 * `capacity_blocks` fail with MSBG_ENOMEM, as try_alloc_block returns None
 * (blockpool.rs:118-131). */
typedef struct msbg_grid msbg_grid;

/* SparseGrid::new -- sparse_grid.rs:107-155 (log2_bsx 0..4: edge 16 >> l of multires level l, grid.rs:57-63). */
int msbg_grid_new(int64_t sx, int64_t sy, int64_t sz, int log2_bsx, float empty_value, float full_value,
                  int64_t capacity_blocks, msbg_grid **out);
int msbg_grid_free(msbg_grid *g);
/* dims[0..7] = sx, sy, sz, nx, ny, nz, n_blocks, n_slots (host). */
int msbg_grid_dims(const msbg_grid *g, int64_t *dims);
/* ensure_blocks_parallel -- sparse_grid.rs:333-355 (host bids, unique). */
int msbg_grid_ensure_blocks(msbg_grid *g, const int64_t *bids, int64_t n);
/* fill_blocks_parallel -- sparse_grid.rs:359-369 (host bids). */
int msbg_grid_fill_blocks(msbg_grid *g, const int64_t *bids, int64_t n, float value);
/* set_full_block / set_empty_block -- sparse_grid.rs:270-282. */
int msbg_grid_set_full_block(msbg_grid *g, int64_t bid);
int msbg_grid_set_empty_block(msbg_grid *g, int64_t bid);
/* set_empty_value / set_full_value -- sparse_grid.rs:253-266. */
int msbg_grid_set_empty_value(msbg_grid *g, float v);
int msbg_grid_set_full_value(msbg_grid *g, float v);
/* Payload upload / download for `bids` (host, n * B^3 floats).  Reading a
 * dummy or absent block returns its sentinel value, as get_value_in_block
 * (sparse_grid.rs:301-308); writing one is MSBG_EINVAL. */
int msbg_grid_write_blocks(msbg_grid *g, const int64_t *bids, int64_t n, const float *host_data);
int msbg_grid_read_blocks(const msbg_grid *g, const int64_t *bids, int64_t n, float *host_out);
/* Bulk load / store of the whole layout (host b2s[n_blocks], slab[n_slots*B^3]). */
int msbg_grid_load(msbg_grid *g, const int32_t *b2s, const float *slab, int64_t n_slots);
int msbg_grid_store(const msbg_grid *g, int32_t *b2s, float *slab, int64_t cap_slots, int64_t *n_slots);
/* The bid -> slot table alone (host b2s[n_blocks]; no device traffic):
 * slot >= 2 is a value block (is_value_block, sparse_grid.rs:199), 0 / 1 the
 * empty / full dummies. */
int msbg_grid_block_slots(const msbg_grid *g, int32_t *b2s);
/* Deep copy of src's layout and values into dst (same dimensions and block
 * size, dst capacity >= src's slots), stream-ordered on the device.  No
 * reference counterpart (SparseGrid has no Clone); it keeps the
 * pre-projection face velocities of a FLIP step. */
int msbg_grid_copy(msbg_grid *dst, const msbg_grid *src);
/* value_block_ptr_mut -- sparse_grid.rs:424-429: device payload, NULL for dummies. */
int msbg_grid_value_block_ptr(msbg_grid *g, int64_t bid, float **dev_ptr);
/* Device views of the slab and the bid->slot table. */
int msbg_grid_device_views(msbg_grid *g, float **slab, int32_t **b2s);

/* ---- colour schedule and sweeps -------------------------------------------- */
/* build_color_buckets -- solver.rs:247-267: the 8 buckets back to back in
 * `out` (host, n entries), sizes in counts[8]; each bucket Morton-sorted. */
int msbg_color_buckets(const msbg_grid *g, const int64_t *active, int64_t n, int64_t *out, int64_t *counts);

/* Sweeper::sweep -- solver.rs:166-217 (with apply_pde, solver.rs:228-240):
 * `iterations` 8-colour in-place passes over `active` (host bids) with the
 * Neumann halo; dummy / absent blocks are skipped.  Results are bit-identical
 * to the reference's arithmetic (fma where it uses mul_add). */
int msbg_sweep(msbg_grid *g, const int64_t *active, int64_t n, int stencil, float dt, int iterations,
               int clamp01);

/* A reusable sweep plan: the colour buckets of `active` resident on the
 * device, run many times without host work (Sweeper::new + repeated sweep). */
typedef struct msbg_sweep_plan msbg_sweep_plan;
int msbg_sweep_plan_new(msbg_grid *g, const int64_t *active, int64_t n, msbg_sweep_plan **out);
int msbg_sweep_plan_free(msbg_sweep_plan *p);
int msbg_sweep_plan_run_async(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01);
/* The same passes replayed as one HIP graph, captured on first use of a
 * stencil / dt / iterations / clamp01 combination.  The plan caches the
 * graphs of its 4 most recently used combinations, so alternating stencils
 * replay without re-capture; an evicted graph is destroyed only after its
 * last launch has completed. */
int msbg_sweep_plan_run_graph_async(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01);
/* The same passes with a HIP event pair around every colour launch: returns
 * the summed kernel time and the launch count (synchronizes). */
int msbg_sweep_plan_run_timed(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01,
                              float *kernel_ms, int64_t *launches);

/* Sweeper<Density> -- solver.rs:154-240 with D = Density (u16): the same
 * sweeps over a u16 slab d_slab16 laid out by `layout`'s bid -> slot table
 * (slots 0 / 1 hold the empty / full values).  The halo is dequantised as
 * u * (1/65535) (Density::to_f32, channel.rs:124-128) and every result is
 * stored as round(v * 65535), half away from zero (StoreBack<Density>,
 * math/stencil.rs:162-174), saturated to [0, 65535]; use clamp01 = 1 for
 * the [0, 1] range the reference's quantiser assumes.  Single level. */
int msbg_sweep_density(msbg_grid *layout, uint16_t *d_slab16, const int64_t *active, int64_t n, int stencil, float dt,
                       int iterations, int clamp01);
/* A plan over a u16 slab (run with msbg_sweep_plan_run_async / _run_graph_async). */
int msbg_sweep_plan_new_density(msbg_grid *layout, uint16_t *d_slab16, const int64_t *active, int64_t n,
                                msbg_sweep_plan **out);

/* Sweeper<Density8> -- solver.rs:154-240 with D = Density8 (u8,
 * channel.rs:55-59,155-189): the same sweeps over a u8 slab d_slab8 laid out
 * by `layout`'s bid -> slot table.  The halo is dequantised linearly as
 * u * (1/255) (the Density8 Dequant, math/sample.rs:64-69: the sqrt-space
 * value) and every result v is stored sqrt-compressed with stochastic
 * rounding (StoreBack<Density8>, math/stencil.rs:176-192): vf = sqrt(v) * 255,
 * code = floor(vf) + (u < vf - floor(vf)), saturated to [0, 255] (v is meant
 * to lie in [0, 1]: clamp01 = 1).  The reference draws u from per-thread
 * lane generators whose stream depends on its SIMD width and schedule; here u
 * is a fixed function of (seed, sweep index, bid, voxel): h0 = fmix32(seed ^
 * sweep * 0x9e3779b9), h = fmix32(fmix32(h0 ^ bid) ^ vid * 0x9e3779b9) (murmur3
 * finaliser; 0 -> 0xdeadbeef), u = float((h * 16807) >> 9 as a [1, 2)
 * mantissa) - 1.  The sweep index counts the sweeps of the call (or of the
 * plan, which keeps counting across runs and graph replays) from 0.
 * Single level. */
int msbg_sweep_density8(msbg_grid *layout, uint8_t *d_slab8, const int64_t *active, int64_t n, int stencil, float dt,
                        int iterations, int clamp01, uint32_t seed);
/* A plan over a u8 slab; its sweep index starts at 0. */
int msbg_sweep_plan_new_density8(msbg_grid *layout, uint8_t *d_slab8, const int64_t *active, int64_t n,
                                 uint32_t seed, msbg_sweep_plan **out);

/* Halo topology -- the neighbour resolution of halo.rs:59-80 (single level,
 * coarse = NULL) or halo.rs:234-251 (multires): for each bid a row of
 * (2H+1)^3 int32 codes, index dx + n*dy + n*n*dz, code >= 0 = the neighbour's
 * slab slot (0 = empty dummy, also outside the domain, 1 = full dummy),
 * code < 0 = a coarser neighbour at coarse slot -code-1.  host_out holds
 * n * (2H+1)^3 entries.  A sweep plan stores these rows. */
int msbg_halo_topology(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *dev_levels, const int64_t *bids,
                       int64_t n, int halo, int32_t *host_out);

/* HaloBlock::fill::<HALO, FULL> -- multires/halo.rs:31-170, batched: writes
 * the HSX^3 tile (HSX = B + 2*halo, index z*HSX^2 + y*HSX + x) of every bid in
 * `bids` (host) to `dev_tiles`.  Faces-only fills write 0 to the x-halo edge
 * and corner cells a fresh HaloBlock holds there (halo.rs:16-20,154). */
int msbg_halo_gather(const msbg_grid *g, const int64_t *bids, int64_t n, int halo, int full, int bc,
                     float *dev_tiles);

/* ---- multires ---------------------------------------------------------------
 * `fine` is level l (edge 2^L), `coarse` level l+1 (edge 2^(L-1)) over the same
 * block grid; `dev_levels[bid]` is the block's level (level0,
 * multires/blockinfo.rs:85). */
/* fill_multires -- multires/halo.rs:187-358, batched like msbg_halo_gather. */
int msbg_halo_gather_multires(const msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                              const int64_t *bids, int64_t n, int halo, int full, int bc, float *dev_tiles);
/* The 8-colour sweep of the level-l blocks in `active` with fill_multires
 * halos (the smoother the multires benches time, multires_bench.rs:143-174). */
int msbg_sweep_multires(msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                        const int64_t *active, int64_t n, int stencil, float dt, int iterations, int clamp01);
int msbg_sweep_plan_new_multires(msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                                 const int64_t *active, int64_t n, msbg_sweep_plan **out);
/* The 8-colour sweep of a multires grid's blocks at every level (extends
 * fill_multires, multires/halo.rs:187-358, whose centre block is always fine,
 * to centre blocks at coarse levels; Sweeper::sweep's colour schedule,
 * solver.rs:166-267).  grids[l] (l < n_levels, 2 <= n_levels <= 5) is level
 * l with block edge 16 >> l over one block grid; dev_levels[bid] is the
 * block's level; each bid of `active` (host) is updated at its own level, in
 * place, and skipped when it has no value block there.  A block's colour is
 * the parity of its block coordinates at any level, and one colour pass
 * updates the active blocks of that colour at all levels.  Halos: a level-0
 * block's is msbg_halo_gather_multires's; a block at level l >= 1 reads each
 * halo voxel g (Neumann mirror of level l's domain) from the level its block
 * lives at: level l itself, level l+1 at g >> 1 (when coarser), or the mean
 * of level l-1's 2^3 voxels 2g + {0,1}^3 summed z, y, x in order and divided
 * by 8 (when finer; downsample_channel_avg's arithmetic,
 * multires/downsample.rs:44-53).  The kernels are the reference's on the
 * level's block edge.  Edge-1 blocks (level 4) take radius-1 stencils only. */
int msbg_sweep_levels(msbg_grid *const *grids, int n_levels, const uint8_t *dev_levels, const int64_t *active,
                      int64_t n, int stencil, float dt, int iterations, int clamp01);
/* The same sweep as a reusable plan (run with msbg_sweep_plan_run_async /
 * _run_graph_async / _run_timed). */
int msbg_sweep_plan_new_levels(msbg_grid *const *grids, int n_levels, const uint8_t *dev_levels,
                               const int64_t *active, int64_t n, msbg_sweep_plan **out);
/* downsample_channel_avg -- multires/downsample.rs:13-58 for `bids` (host);
 * the coarse blocks are ensured first. */
int msbg_downsample_avg(msbg_grid *coarse, const msbg_grid *fine, const int64_t *bids, int64_t n);
/* regularize_refinement_map -- multires/refinement.rs:66-107 on a device
 * level map of nx*ny*nz bytes, in place. */
int msbg_regularize_levels(uint8_t *dev_levels, int64_t nx, int64_t ny, int64_t nz, int n_levels);
/* compute_block_topology -- multires/refinement.rs:124-208 (levelMg 0, no
 * obstacles) on a device level map of nx*ny*nz bytes: d_flags_out[b] (device,
 * u16) = d_user_flags[b] (device, or NULL for none) | BLK_DOM_BORDER (1 << 1)
 * on the block-grid faces | BLK_FINE_COARSE (1 << 3) if a 26-neighbour is
 * coarser | BLK_COARSE_FINE (1 << 2) if one is finer (blockinfo.rs:39-54).
 * d_fc_bids_out (device, room for nx*ny*nz) receives Topology::
 * blocks_fine_coarse, level 0 first, ascending within a level;
 * fc_counts[0..n_levels) (host) their lengths; *n_act (host) =
 * Topology::n_act_blocks, the BLK_EXISTS (1 << 12) count.  The reference's
 * store.level0 is the level map itself. */
int msbg_block_topology(const uint8_t *d_levels, int64_t nx, int64_t ny, int64_t nz, int n_levels,
                        const uint16_t *d_user_flags, uint16_t *d_flags_out, int64_t *d_fc_bids_out,
                        int64_t *fc_counts, int64_t *n_act);

/* ---- particles ---------------------------------------------------------------
 * Particle arrays are device pointers: positions and velocities are n x 3
 * float (x, y, z interleaved), one particle per row, as fluid/mod.rs:30-36
 * stores them.  Counts and block ids are below 2^31. */

/* morton3 -- multires/sort.rs:20-22 -- of each particle's voxel (truncated
 * position, as p2g.rs:153), low 30 bits: exact for voxel coordinates < 1024. */
int msbg_morton_keys(const float *d_pos, int64_t n, uint32_t *d_keys);
/* Stable LSD radix sort of (key, value) pairs by the low key_bits bits, in
 * place (8-bit digits; wave ballot ranking, no global atomics).  The
 * reference has no radix sort: it is the permutation std::stable_sort gives. */
int msbg_radix_sort_pairs(uint32_t *d_keys, uint32_t *d_vals, int64_t n, int key_bits);
/* Morton sort of particles: d_keys = sorted 30-bit keys, d_perm[i] = original
 * index of the i-th particle (stable: ties keep input order). */
int msbg_morton_sort(const float *d_pos, int64_t n, uint32_t *d_keys, uint32_t *d_perm);
/* Row-major block id of each particle -- p2g.rs:149-157. */
int msbg_particle_bids(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, int32_t *d_bids);
/* bucket_indices -- particles/sort.rs:174-272: d_perm (n) = block-major
 * permutation, stable by block id (the order of the reference's dense oracle
 * bucket_by_block_dense, particles/sort.rs:455-490; the reference's own
 * intra-block order is unspecified, :234-237), d_particle_blocks (<= n)
 * ascending occupied blocks, d_starts (n_occupied + 1) CSR offsets. */
int msbg_bucket_indices(const int32_t *d_bids, int64_t n, int64_t n_blocks, int32_t *d_perm,
                        int32_t *d_particle_blocks, int32_t *d_starts, int64_t *n_occupied);
/* active_blocks -- particles/active.rs:15-47 with footprint_axis
 * (particles/mod.rs:207-212): union of every particle's footprint blocks
 * trunc(p/16 -+ r_scan/16) clipped to the block grid, ascending (d_out holds
 * up to nx*ny*nz entries).  P2G uses r_scan = r_p + 1 (p2g.rs:104-107). */
int msbg_active_blocks(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, float r_scan,
                       int32_t *d_out, int64_t *n_out);
/* No counterpart in the reference.  msbg_active_blocks for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): same arguments, same results, and
 * no hipMalloc / hipFree / allocating library call.  Its scratch (block
 * flags, ranks and scan partials) is one fixed device array of the library
 * per GPU (2.1 MB, held until the process ends), so the block grid is limited
 * to MSBG_NOALLOC_MAX_BLOCKS blocks (a 1024^3 domain of 16^3 blocks); larger
 * grids return MSBG_EINVAL before any device work.  Every call on a GPU shares
 * that array: calls that overlap on one GPU (several host threads, or streams
 * set in between) race.  Synchronizes the library stream (the count is read
 * back). */
#define MSBG_NOALLOC_MAX_BLOCKS 262144
int msbg_active_blocks_noalloc(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, float r_scan,
                               int32_t *d_out, int64_t *n_out);
/* d_dst[i] = d_src[d_perm[i]] for rows of row_bytes (a multiple of 4): the
 * payload gather of bucket_particles (fluid/mod.rs:84-89). */
int msbg_gather_rows(const void *d_src, const uint32_t *d_perm, int64_t n, int64_t row_bytes, void *d_dst);
/* msbg_morton_sort followed by the payload gathers of positions (n x 3),
 * velocities (n x 3) and masses (n; d_mass null: all 1) into the *_out
 * arrays in sorted order -- the particle set bucket_by_block returns
 * (particles/sort.rs:274-285) with the payload gather of fluid/mod.rs:84-89.
 * Same keys, permutation and gathered values as msbg_morton_sort plus three
 * msbg_gather_rows calls; the payload travels as one 32-byte record per
 * particle, written while the keys are computed. */
int msbg_morton_sort_particles(const float *d_pos, const float *d_vel, const float *d_mass, int64_t n,
                               uint32_t *d_keys, uint32_t *d_perm, float *d_pos_out, float *d_vel_out,
                               float *d_mass_out);

/* ---- grid <-> particle transfers --------------------------------------------- */
/* Sampler::sample -- math/sample.rs:489-526, with gather_map
 * (math/gather.rs:54-158): d_out[i] = sample at d_pos[i]; interp 0 = Linear,
 * 1 = CubicBSpline; align 0 = Corner, 1 = CellCentered; bc as above. */
int msbg_sample(const msbg_grid *g, const float *d_pos, int64_t n, int interp, int align, int bc, float *d_out);
/* MacSampler::sample -- fluid/g2p.rs:26-32: face a sampled cell-centred at
 * p - 1/2 e_a; d_out is n x 3. */
int msbg_mac_sample(msbg_grid *const *vel, const float *d_pos, int64_t n, int interp, int bc, float *d_out);
/* g2p_blend -- fluid/g2p.rs:42-45 over n x 3 velocities. */
int msbg_g2p_blend(const float *d_u_old, const float *d_u, const float *d_u_tilde, int64_t n, float alpha,
                   float *d_out);
/* The G2P of a FLIP step -- fluid/g2p.rs:11-45 composed per particle:
 * u = MacSampler(vel).sample(p), u* = MacSampler(vel_tilde).sample(p),
 * out = g2p_blend(u_old, u, u*, alpha).  Same arithmetic (and bits) as two
 * msbg_mac_sample calls and msbg_g2p_blend, in one pass over the particles. */
int msbg_g2p_flip(msbg_grid *const *vel, msbg_grid *const *vel_tilde, const float *d_pos, const float *d_u_old,
                  int64_t n, int interp, int bc, float alpha, float *d_out);
/* particles_to_grid -- fluid/p2g.rs:141-164: bucket by block, active blocks
 * (footprint r_p + 1) ensured and zeroed on all six channels
 * (MacGrid::zero_blocks, mac.rs:95-102), the cubic-falloff splat of mass and
 * momentum to faces (p2g.rs:169-412) in 8 colour passes per direction with
 * LDS staging, then u = P/M (mac.rs:106-132).  vel[3] and mass[3] are the face
 * channels x, y, z over one domain; 0 < r_p <= 3.  The commit adds each margin
 * region over its own width (the semantics of the reference's naive oracle,
 * p2g.rs:445-502).  Sums within a staging cell are order-dependent in the last
 * bits, as in the reference (particles/sort.rs:234-237). */
int msbg_particles_to_grid_mac(msbg_grid *const *vel, msbg_grid *const *mass, const float *d_pos, const float *d_vel,
                               const float *d_mass, int64_t n, float r_p, int64_t *n_active);
/* Trilinear (order 1) or quadratic B-spline (order 2) P2G onto a cell-centred
 * grid (no reference counterpart): blocks under the particles' r_scan = 2
 * footprint are ensured and zeroed on all four grids; particles are Morton
 * sorted and every cell gathers its 27 neighbour voxels in a fixed order, so
 * the result is deterministic.  mass[cell] = sum w m, vel[a][cell] =
 * sum w m u_a / mass (0 where mass is 0).  d_mass may be NULL (m = 1).
 * Particles lie in [0, s) per axis, s <= 1024. */
int msbg_p2g_bspline(int order, msbg_grid *mass, msbg_grid *const *vel, const float *d_pos, const float *d_vel,
                     const float *d_mass, int64_t n, int64_t *n_active);
/* The matching G2P: d_out (n x 3) = sum over the support cells of w * vel. */
int msbg_g2p_bspline(int order, msbg_grid *const *vel, const float *d_pos, int64_t n, float *d_out);

/* MacGrid::divergence -- fluid/mac.rs:137-171: ensures `active` (host bids)
 * in dst and writes ((u(x) - u(x-1)) + (v(y) - v(y-1))) + (w(z) - w(z-1)) per
 * cell, a face at index 0 on its axis reading 0.  vel[3] and dst share the
 * domain.  Each cell is stored at its layout index (the reference's running
 * index, mac.rs:155-166, coincides with it for whole 16^3 blocks). */
int msbg_mac_divergence(msbg_grid *const *vel, msbg_grid *dst, const int64_t *active, int64_t n);
/* Pressure projection of the face velocities.  No reference counterpart: the
 * reference stops at the divergence (mac.rs:134-171; its projection is the
 * planned step 18).  For every cell c of the listed blocks (host bids) and
 * axis a, the face u_a(c) between c and c + e_a (the layout
 * msbg_mac_divergence reads) becomes u_a(c) - scale (p(c + e_a) - p(c)),
 * the product and difference rounded separately.  p is a field over
 * p_layout's slots (a PoissonBand field): a neighbour outside the domain is
 * the cell itself (zero gradient), one in an absent block reads slot 0.  With
 * A p = div (the band's operator) and scale = -1 the projected field's
 * divergence equals div - A p, the solve's residual. */
int msbg_mac_subtract_gradient(msbg_grid *const *vel, const msbg_grid *p_layout, const float *d_p,
                               const int64_t *active, int64_t n, float scale);
/* The same projection from a compensated pressure p + p_lo (the x, x_lo pair
 * of a msbg_poisson_solve with x_lo): the gradient is (p(c + e_a) - p(c)) +
 * (p_lo(c + e_a) - p_lo(c)).  No reference counterpart. */
int msbg_mac_subtract_gradient_pair(msbg_grid *const *vel, const msbg_grid *p_layout, const float *d_p,
                                    const float *d_p_lo, const int64_t *active, int64_t n, float scale);
/* particles::splat -- particles/splat.rs:92-319 with D = Density (u16,
 * channel.rs:303-375): the active blocks of the particles (footprint r_scan,
 * active.rs:15-47) are ensured in the layout grid g and filled with 65535 in
 * the u16 slab d_slab16 (slab_slots * 4096 elements over g's bid -> slot
 * table; slot 0 / 1 are set to 0 / 65535, the grid's Density(0) /
 * Density(u16::MAX) dummies); then every voxel within r_scan of a particle,
 * in the 16 + 2 ceil(r_scan) window of the particle's block, takes
 * min(voxel, round_half_away(d^2 / r_scan^2 * 65535)), d measured to the
 * voxel centre.  A window region (interior / margin slab, splat.rs:263-313)
 * is committed only when it lies inside the domain and targets a value block;
 * margin regions are committed over their own width (the reference's 4-lane
 * margin rows coincide with it for 3 < r_scan <= 4 and spill past it below).
 * Particles lie in [0, s) per axis; 0 < r_scan <= 6.  *n_active = the number
 * of active blocks. */
int msbg_density_splat(msbg_grid *g, uint16_t *d_slab16, int64_t slab_slots, const float *d_pos, int64_t n,
                       float r_scan, int64_t *n_active);

/* ---- surface-reconstruction tail (particles/mod.rs:214-302) ------------------
 * place -- particles/sort.rs:31-143: instance i (0 <= i < n_instances; 0 = one
 * per base point, n_instances <= n_base) of the base cloud d_base (device,
 * n_base x 3) sits at 0.2 smax + 0.6 smax (b_i - bbox_min) / span_max and
 * holds every base point j at origin + smin instance_scale_factor (b_j -
 * bbox_min) / span_max (smin / smax: the smallest / largest domain edge, f32
 * operations in the reference's order, none fused).  A point is kept when it
 * lies within DomainBounds (align24(0.005) .. align24(size - 0.005),
 * particles/mod.rs:180-199) and its truncation lies in the grid.  Kept
 * positions (device, n x 3) and centre block ids (device) are written in
 * instance-major, base-point order; *n_out = their count (MSBG_ENOMEM when it
 * exceeds `capacity`, *n_out still set; with both outputs NULL and capacity 0
 * only the count is returned).  Placed::active is the footprint
 * union of the kept positions: msbg_active_blocks with r_scan. */
int msbg_surface_place(const float *d_base, int64_t n_base, const float *bbox_min, float span_max, int64_t sx,
                       int64_t sy, int64_t sz, int64_t n_instances, float instance_scale_factor, float *d_pos_out,
                       int32_t *d_bids_out, int64_t capacity, int64_t *n_out);
/* finalize -- particles/finalize.rs:35-71 on the Density (u16) slab of a
 * layout grid, for the `active` (host) blocks that have a value slot:
 * ratio = u / 65535 (times the f32 constant 1/65535); touched = ratio < 0.999;
 * f = 1 - clamp((sqrt(ratio r_scan^2) - r_p + r_p) / (2 r_p + nb_dist), 0, 1)
 * with r_scan = r_p + nb_dist; stored as round(f * 65535) (half away from
 * zero), 0 where untouched. */
int msbg_density_finalize(const msbg_grid *layout, uint16_t *d_slab16, int64_t slab_slots, const int64_t *active,
                          int64_t n, float r_particle, float nb_dist);

/* ---- Poisson relaxation on a band of value blocks ------------------------
 * No reference counterpart: the reference's solver is planned, not built
 * (docs/roadmap.md:610-702; the weighted Jacobi formula y = omega (b + sum nb)
 * D + (1 - omega) x is planned at docs/roadmap.md:622-626).  The unknowns are
 * the cells of the listed 16^3 value blocks of a layout grid; every field is a
 * device slab shaped like the layout grid's (n_slots * 4096 elements of the
 * storage type, slot-major, slots 0 / 1 = values read for absent / full
 * neighbour blocks).  Operator: (A x)_c = 6 x_c - sum of the 6 face
 * neighbours, a neighbour outside the domain being the cell itself (Neumann
 * mirror, math/boundary.rs:35-56).  One relaxation of cell c:
 *   s = ((((x-x + x+x) + x-y) + x+y) + x-z) + x+z,  t = fmaf(-6, x_c, b_c + s),
 *   x_c' = fmaf(omega / 6, t, x_c) rounded to the storage type (RNE).
 * Storage types: FP32, IEEE binary16, bfloat16; arithmetic is FP32. */
enum { MSBG_STORAGE_F32 = 0, MSBG_STORAGE_F16 = 1, MSBG_STORAGE_BF16 = 2 };
enum { MSBG_SOLVE_DIRECT = 1 };
typedef struct msbg_poisson msbg_poisson;
typedef struct msbg_learned msbg_learned;
typedef struct {
    int storage;          /* MSBG_STORAGE_*: storage the sweeps read and write */
    int refine;           /* 0: red-black sweeps on x itself (storage must be F32);
                             1: FP32 iterative refinement: r = b - A x (FP32), rt = r / 2^k
                                (2^k ~ rms(r)), e = 0, inner sweeps on A e = rt in `storage`,
                                x = fmaf(2^k, e, x);
                             2: the same refinement step with one multigrid V-cycle as the inner
                                solve: block-aligned cell-centred levels (16^3, 8^3, 4^3, 2^3 cells
                                per block, one per block, then the block lattice coarsened by 2),
                                red-black smoothing (`storage` on the band, FP32 below), trilinear
                                prolongation, FP64 conjugate gradients on the coarsest level
                                (<= 1024 cells); sweeps_per_check = smoothing sweeps before and
                                after the coarse correction on every level; the solve also stops
                                (converged = 0) after two cycles in a row that each take off less
                                than 10% of the residual (x at its FP32 rounding floor) */
    float omega;          /* SOR factor, 0 < omega < 2 */
    int sweeps_per_check; /* red-black sweeps between FP64 residual checks (inner sweeps per refinement step) */
    int max_sweeps;       /* cap on the band's red-black sweeps (refine = 2: 2 sweeps_per_check per cycle) */
    int flags;            /* 0, or MSBG_SOLVE_DIRECT: launch every check period's kernels directly on the
                             library stream instead of replaying a captured HIP graph (for callers that
                             capture the solve themselves, and for profilers that trace graph launches
                             poorly); same kernels, same results */
    double tol;           /* stop when ||b - A x||_2 <= tol * ||b||_2 (FP64 norms) */
    msbg_learned *learned;   /* NULL, or a network proposing a correction of x after every check period */
    const msbg_grid *phi;    /* phase field (network input; layout dimensions, 16^3 blocks) */
    float *x_lo;          /* NULL, or (refine = 1 / 2, no learned correction) an FP32 field holding the low
                             part of a compensated iterate x + x_lo (zero, or a previous solve's low part, on
                             entry): the refinement updates the pair with error-free FP32 sums and computes
                             r = b - A (x + x_lo) compensated, so the residual can fall below the floor one
                             FP32 field sets (~2.3e-7 ||x|| / ||b||); on return x = round(x + x_lo) */
} msbg_poisson_params;
typedef struct {
    int sweeps;           /* red-black sweeps done */
    int checks;           /* residual evaluations */
    int converged;        /* 1 when the tolerance was met */
    int corrections;      /* learned corrections tried */
    int accepted;         /* ... and accepted (the true residual decreased) */
    int levels;           /* refine = 2: multigrid levels including the band; else 0 */
    double rel_residual;  /* ||b - A x|| / ||b|| at the last check */
    double r2, b2;        /* squared norms at the last check */
    float ms;             /* wall time of the solve on the library stream (HIP events) */
} msbg_poisson_stats;

int msbg_poisson_new(const msbg_grid *layout, const int64_t *active_blocks, int64_t n_active, msbg_poisson **out);
int msbg_poisson_free(msbg_poisson *p);
int msbg_poisson_info(const msbg_poisson *p, int64_t *n_active, int64_t *n_cells, int64_t *n_slots);
/* One weighted Jacobi sweep x_out = x_in + omega/6 (b - A x_in) over the listed
 * cells (x_in != x_out; other cells of x_out are not written). */
int msbg_poisson_jacobi(msbg_poisson *p, int storage, const void *x_in, void *x_out, const void *b, float omega);
/* `sweeps` (even) weighted Jacobi sweeps x -> tmp -> x ..., replayed as one
 * HIP graph; tmp must hold x's values outside the listed cells (e.g. a copy
 * of x), the result is in x. */
int msbg_poisson_jacobi_sweeps(msbg_poisson *p, int storage, void *x, void *tmp, const void *b, float omega,
                               int sweeps);
/* One red-black SOR half sweep in place over the listed cells with
 * (x + y + z) & 1 == color. */
int msbg_poisson_rbgs(msbg_poisson *p, int storage, void *x, const void *b, int color, float omega);
/* r = b - A x over the listed cells (r: FP32 slab or NULL) and the FP64 squared
 * norms of r and b (synchronous). */
int msbg_poisson_residual(msbg_poisson *p, int storage, const void *x, const void *b, float *r, double *r2,
                          double *b2);
/* r = b - A (x + x_lo) of a compensated FP32 iterate (see x_lo below), rounded
 * once to FP32 (r: slab or NULL), and the FP64 squared norms of r and b
 * (synchronous).  No reference counterpart. */
int msbg_poisson_residual_pair(msbg_poisson *p, const float *x, const float *x_lo, const float *b, float *r,
                               double *r2, double *b2);
/* Element-wise storage conversion of n_elems values (RNE on narrowing). */
int msbg_poisson_convert(int src_storage, const void *src, int dst_storage, void *dst, int64_t n_elems);
/* Solve A x = b (x, b: FP32 slabs; x holds the initial guess) with red-black
 * SOR, checking the FP64 relative residual every sweeps_per_check sweeps.  Each
 * check period runs as one replayed HIP graph. */
int msbg_poisson_solve(msbg_poisson *p, const msbg_poisson_params *params, float *x, const float *b,
                       msbg_poisson_stats *stats);

/* ---- learned error correction ---------------------------------------------
 * No reference counterpart (the reference has no learned component).  A
 * 4-layer 3^3 convolution network per 16^3 block: input X0 = 20^3 x 2 bf16
 * (the block with a 2-cell halo; channel 0 = residual / 2^k, channel 1 =
 * phase field), L1 valid 2->16, L2 valid 16->16, L3 same 16->16 (ReLU after
 * each), L4 same 16->1, output Y = 16^3 f32 (the error estimate / 2^k /
 * out_scale).  Tensors are channels-last [z][y][x][c].  Parameters (FP32):
 * 4 layers x (W[27][16][16] + bias[16]), tap t = dx + 3 dy + 9 dz; layer 1
 * uses input channels 0..1, layer 4 output channel 0, the rest is zero.
 * Forward and backward run as BF16 MFMA implicit GEMMs with FP32 accumulation. */
int msbg_learned_param_count(int *n);
int msbg_learned_new(msbg_learned **out); /* parameters zero */
int msbg_learned_free(msbg_learned *net);
int msbg_learned_set_params(msbg_learned *net, const float *host_params, float out_scale); /* resets Adam */
int msbg_learned_get_params(const msbg_learned *net, float *host_params, float *out_scale);
/* d_y[t] (16^3 f32) = network output for d_x0[t] (20^3 x 2 bf16), t < n. */
int msbg_learned_forward(msbg_learned *net, const uint16_t *d_x0, int64_t n, float *d_y);
/* Loss sum (y - target)^2 / (2 n 4096) over n tiles and its gradient with
 * respect to the parameters (host_grad: msbg_learned_param_count floats). */
int msbg_learned_loss_grad(msbg_learned *net, const uint16_t *d_x0, const float *d_target, int64_t n,
                           double *loss, float *host_grad);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * pairs drawn uniformly (SplitMix64 from `seed`) from n_pairs (x0, target)
 * pairs; host_losses[step] = the batch loss before the step. */
int msbg_learned_train(msbg_learned *net, const uint16_t *d_x0, const float *d_target, int64_t n_pairs, int steps,
                       int batch, float lr, uint64_t seed, float *host_losses);
/* Network input of every listed block of a Poisson band (d_x0: n_active x
 * 20^3 x 2): channel 0 = d_r * inv_scale (0 in unlisted blocks), channel 1 =
 * the phase field; halo cells outside the domain mirror. */
int msbg_poisson_net_input(msbg_poisson *p, const msbg_grid *phi, const float *d_r, float inv_scale,
                           uint16_t *d_x0);
/* Training target of every listed block: (x* - x) * inv_scale (n_active x 16^3). */
int msbg_poisson_error_target(msbg_poisson *p, const float *d_x, const float *d_xstar, float inv_scale,
                              float *d_target);
/* One learned correction of x: r = b - A x, 2^k ~ rms(r), delta = 2^k
 * out_scale Y(net input), x' = x + alpha delta with alpha = (r . A delta) /
 * |A delta|^2 (FP64), kept only if ||b - A x'|| < ||b - A x|| (the true
 * residual decides; otherwise x is unchanged). */
int msbg_poisson_learned_correct(msbg_poisson *p, msbg_learned *net, const msbg_grid *phi, float *x, const float *b,
                                 double *r2_before, double *r2_after, int *accepted);

#ifdef __cplusplus
}
#endif
#endif /* MSBG_HIP_H */
