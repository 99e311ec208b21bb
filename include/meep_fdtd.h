/*
 * meep_fdtd.h -- C ABI of the MI355X (gfx950) FDTD time-stepping engine.
 *
 * Plain pointers and sizes only. Every volume crosses the boundary in the engine's device
 * layout, x fastest: element (i, j, k) of an (nx, ny, nz) grid is at ((k*ny)+j)*nx+i. (The
 * reference stores C-order (nx, ny, nz), z fastest; the Python host transposes on the way
 * in and out.) Every call returns FDTD_OK or an error code; fdtd_last_error() gives the
 * text of the calling thread's last failure. Nothing falls back to another path.
 *
 * Each entry point names the reference interface it replaces (file:line in the reference
 * meep_gpu package). The reference reaches its kernels through one seam,
 * FastPathPlan.dispatch(slot, fields) (fastpath.py:4631) over DRIVER_SLOTS (fastpath.py:575)
 * = step_B, fill_B, update_H, step_D, fill_D, update_E, update_P; the Python plan object in
 * meep_gpu.fastpath answers those slots by calling fdtd_step_h / fdtd_step_d below.
 */
#ifndef MEEP_FDTD_H
#define MEEP_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FDTD_OK = 0,
    FDTD_EINVAL = 1,  /* bad argument or shape */
    FDTD_EHIP = 2,    /* a HIP runtime call failed */
    FDTD_ESTATE = 3,  /* call not valid in the simulation's current state */
    FDTD_ENOMEM = 4   /* device allocation failed */
};

/* Lorentz/Drude poles one simulation holds (up to 4 step in the lossless kernels; more select
 * the general ones, as a conductivity or a metallic wall does). */
#define FDTD_MAX_POLES 8

typedef struct fdtd_sim fdtd_sim;
typedef struct fdtd_batch fdtd_batch;

typedef struct {
    int32_t n[3];         /* stored cells per axis: nx, ny, nz (grid.py:404-592) */
    int32_t invariant[3]; /* 1 for a translationally invariant one-cell axis (2-D runs) */
    float dtdx;           /* dt/dx as float32 (stepping.py:361) */
    int32_t pml;          /* 1 when an absorber is installed (fields.py:708) */
    double dt;            /* the timestep, for monitor sample times (grid.py:428) */
} fdtd_config;

/* Text of the calling thread's last error (empty string when none). */
const char *fdtd_last_error(void);

/* Engine build and device description, for reports. */
int fdtd_device_info(char *text, size_t size);

/* Free and total device memory of the current GPU (hipMemGetInfo), for footprint admission of
 * served solves. No counterpart in the reference (it runs one simulation per process). */
int fdtd_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

/* Create a simulation: D, B, E always; H, fu_B, fu_D, f_w_H, f_w_E when cfg->pml.
 * Replaces Fields allocation (fields.py:605-760) and enable_pml_storage (fields.py:708).
 * All volumes start at +0; inverse epsilon starts at 1. */
int fdtd_sim_create(const fdtd_config *cfg, fdtd_sim **out);
/* Release a simulation (waits for its stream). A member of a live batch is not freed at once:
 * its release is deferred to fdtd_batch_destroy (fdtd_last_error says so). */
void fdtd_sim_destroy(fdtd_sim *sim);

/* Bytes of device memory the simulation holds. */
int fdtd_sim_bytes(const fdtd_sim *sim, uint64_t *bytes);

/* One axis of the UPML table and its chunk membership line: half 0 = integer lattice,
 * 1 = half-integer; vectors of length n[axis]. The line is whatever the chunk division says
 * (the one-rank rule, or a lifted division's ChunkLayout.membership; at most 2 runs of
 * member rows); on a metallic axis the unowned wall row reads 0. Replaces
 * PML._compute_coefficients (pml.py:673) and RegionPlan membership (chunk_map.py:180, :534,
 * :1093). */
int fdtd_set_layer(fdtd_sim *sim, int32_t half, int32_t axis, const float *kms,
                   const float *sinv, const float *kps, const uint8_t *member);

/* Diagonal chi1inv for E component c (0..2), one volume (fields.py:1396, :1523). */
int fdtd_set_inverse_epsilon(fdtd_sim *sim, int32_t component, const float *volume);

/* Bytes per cell the D pair streams for each component's inverse epsilon: 1 or 2 when the
 * volume has at most 256 or 65536 distinct words (stored as indices into a table of them, the
 * same words), else 4. No counterpart in the reference (its chi1inv is a float32 volume). */
int fdtd_inverse_epsilon_bytes(fdtd_sim *sim, int32_t out[3]);

/* Cells (of nx*ny*nz, in `total`) on which the D pair runs pole work: per row, the x range
 * covering every cell where some pole's sigma or a stored polarization is nonzero (elsewhere
 * every P stays +0 and D - sum P is D). No counterpart in the reference (it updates P over the
 * whole grid, dispersion.py:689). */
int fdtd_pole_rows(fdtd_sim *sim, int64_t *active, int64_t *total);

/* How the fused launches split the grid for runs of `vec` (1, 2, 4) cells: out[0..5] = the
 * box x0, x1, y0, y1, z0, z1 (whole padded rows, x0 = 0 and x1 = the row pitch, whose y and z are
 * on no layer line: only the x line's forms hold there), out[6] = its runs, out[7] = the runs of
 * the shell (the rows on a y or z layer line, each run holding a real cell). No counterpart in
 * the reference (its region plan launches one kernel per region, chunk_map.py:1093). */
int fdtd_work_split(fdtd_sim *sim, int32_t vec, int32_t out[8]);

/* Off-diagonal chi1inv[row][partner] volume (fields.py:1396, stepping.py:1594). A row with
 * any partner switches update_E to the double-precision row form. */
int fdtd_set_offdiagonal(fdtd_sim *sim, int32_t row, int32_t partner, const float *volume);

/* One Lorentz/Drude pole (dispersion.py:243, :645). consts = {2-omega0dtsqr_denom, gamma1,
 * omega0dtsqr, gamma1inv}; sigma[c] a volume, or NULL with sigma_scalar[c] a uniform value;
 * driven[c] 0 leaves component c without a P (a trivial sigma). Returns the pole index. */
int fdtd_add_pole(fdtd_sim *sim, const float consts[4], const float *const sigma[3],
                  const float sigma_scalar[3], const int32_t driven[3], int32_t *pole);

/* A source deposit over the box [lo, lo+cnt): amplitudes amp_re/amp_im per box cell (x
 * fastest), and the complex current J for steps 0..nsteps-1 as (re, im) pairs: the electric
 * table at (n+1/2)*dt, the magnetic at n*dt. field 0 deposits into D (after step_D), 1 into
 * B (after step_B). Replaces sources._deposit (sources.py:1357) + host_writes.subtract_double
 * (host_writes.py:140): f = float(double(f) - (amp_re*J_re - amp_im*J_im)*dt), times the
 * cell's condinv word on a conductive component.
 * field 2 is an electric polarization (MEEP is_integrated) source, which replaces
 * sources.f_minus_p_begin/_end (sources.py:1697-1760): nothing is deposited into D; the table
 * is the dipole P at (n+1)*dt, and update_E reads float(double(D) - (amp_re*P_re -
 * amp_im*P_im)) at the box's cells (no dt, no condinv) while D keeps its word; an off-diagonal
 * chi1inv row reads the same replaced words at its neighbouring cells. field 3 (a magnetic
 * polarization source) returns FDTD_EINVAL.
 * The subtraction is at the box's own cells only: on a folded grid the caller passes a box
 * without its ghost rows plus its images on them (amplitudes times the fold parity), and none
 * on a metallic wall row, as MEEP's boundary passes over f_minus_p leave them. */
int fdtd_add_source(fdtd_sim *sim, int32_t field, int32_t component, const int32_t lo[3],
                    const int32_t cnt[3], const double *amp_re, const double *amp_im,
                    const double *current, int32_t nsteps, double dt);

/* Replaces source `index`'s current table (the order fdtd_add_source added them) with `nsteps`
 * (re, im) float64 pairs from step 0: a run longer than the table first planned extends it
 * here instead of re-creating the simulation. The reference evaluates src_time::current on the
 * host at every step (sources.py:1844-1936, inject :2168-2173); this engine holds the values
 * for the steps ahead. Refused while the simulation is stepped by a batch. */
int fdtd_set_source_currents(fdtd_sim *sim, int32_t index, const double *current, int32_t nsteps);

/* Upload / download one volume by name: "Dx".."Bz", "Ex".., "Hx".., "fu_Bx".., "fu_Dx"..,
 * "f_w_Ex".., "f_w_Hx".., "P[k].Ex", "P_prev[k].Ex" (the reference's digest names,
 * array_path_oracle.py:709). count = nx*ny*nz floats, x fastest. "Hx".. exist only with a
 * PML (without one H is B's words: pass "Bx"..). `host` may be pageable: the copy is staged
 * through the library's pinned buffer and has completed when the call returns. */
int fdtd_upload(fdtd_sim *sim, const char *name, const float *host, size_t count);
int fdtd_download(fdtd_sim *sim, const char *name, float *host, size_t count);

/* The step counter (driver.py:3592) and MEEP's one-time H start (stepping.py:2789). */
int fdtd_get_step(fdtd_sim *sim, int32_t *step);
int fdtd_set_step(fdtd_sim *sim, int32_t step, int32_t h_separated);

/* E = chi1inv*(D - sum P) everywhere, for a run whose E the reference does not store (no
 * absorber, pole or off-diagonal row): its get_E (fields.py:1051). Call once after upload. */
int fdtd_prime_e(fdtd_sim *sim);

/* The fused H pair: step_B + magnetic deposit + update_H (stepping.py:333, :1216;
 * driver.py:3533-3556). Answers the step_B and update_H slots. */
int fdtd_step_h(fdtd_sim *sim);

/* The fused D pair: step_D + electric deposit + update_E + update_P, then the step counter
 * advances (stepping.py:463, :1243, :1962; driver.py:3560-3592). Answers step_D, update_E,
 * update_P. */
int fdtd_step_d(fdtd_sim *sim);

/* nsteps whole steps (H pair, D pair, flux accumulation) on the simulation's stream. */
int fdtd_run(fdtd_sim *sim, int32_t nsteps);

/* Wait for the simulation's stream. */
int fdtd_synchronize(fdtd_sim *sim);

/* Run the float64 transform pass over every plane's buffered samples now, on the simulation's
 * stream (reads do this for their plane; a timed run calls it so that all the samples of its
 * window are transformed inside it). The counterpart of fdtd_batch_flush for a lone
 * simulation; the reference transforms at every step (dft.py:3031), so it has no call for it. */
int fdtd_flush(fdtd_sim *sim);

/* A planar flux monitor over the box [lo, lo+cnt) (one count is 1 along normal_axis):
 * nf angular frequencies omega, scale = dt*decimation/sqrt(2*pi). The four tangential
 * components are sampled at cell centres each accumulating step and transformed in float64
 * (phase exp(i*omega*t) evaluated in double per step; dft.py:316, :3031). Returns its id. */
int fdtd_add_flux_plane(fdtd_sim *sim, int32_t normal_axis, const int32_t lo[3],
                        const int32_t cnt[3], int32_t nf, const double *omega, double scale,
                        int32_t decimation, int32_t *plane);

/* The boundary of one axis: kind 0 periodic (the default), 1 metallic -- a perfect electric
 * conductor on both faces: curl operands past a wall read zero, and the samples on a wall
 * (stored cell 0 of a component whose Yee shift there is 0: tangential D, normal B) are not
 * stepped and are held at zero after the sub-step's sources. Replaces the reference's METALLIC
 * ghost rule and zero_metal_D / zero_metal_B (stepping.py:140-146, :2429-2626, :3094-3183).
 * A layer on the axis (MEEP's PML inside metallic walls) steps its forms with the wall cells'
 * curl masked to zero (stepping.py:2589-2626); an off-diagonal chi1inv row reads zero past the
 * wall and drops its coupling on the wall planes (stepping.py:1594-1745). Refused on an
 * invariant axis. */
int fdtd_set_boundary(fdtd_sim *sim, int32_t axis, int32_t kind);

/* A mirror fold of one axis (the grid then stores MEEP's halved count N - N/2 + 1, plus one
 * over periodic boundaries): kind 0 none, 1 over metallic boundaries, 2 over periodic ones;
 * phase +1 (even mirror) or -1 (odd); far_row (kind 2) the stored row the last slot images,
 * N - stored + 2. Cell 0 of a component whose Yee shift on the axis is 0 and, for kind 2, the
 * last slot of a shift-1 one are ghosts: their curl is dropped and, after the sub-step's
 * sources, each takes its image times MEEP's parity (vector components odd about the plane
 * normal to their own axis, pseudovectors even; times the phase), H updated from the new B on
 * the layer rows. Past the top of a kind-1 axis a curl operand reads zero. Replaces the
 * reference's MIRROR ghost rules, _mask_non_owned_cells and the fill passes
 * fill_symmetry_bc_D/_B and fill_folded_far_ghosts_D/_B (stepping.py:2004-2160, :2361-2626;
 * grid.py:404-575). Refused on an invariant axis, with flux or DFT regions and with
 * polarization (integrated) sources. */
int fdtd_set_fold(fdtd_sim *sim, int32_t axis, int32_t kind, int32_t phase, int32_t far_row);

/* Complex field storage as MEEP keeps it (two realnum parts, stepped with the same real
 * coefficients; reference part_products.py:1-40): `re` and `im` are two simulations set up
 * alike (grid, layer, materials, poles, folds), holding the real and the imaginary parts of
 * every stored volume; a source's imaginary line is the real line of amplitudes (amp_im,
 * -amp_re). kinds[axis] is the Bloch connection of a wrapping axis (0 none, 1 a copy, 2 a
 * negation, 3 a product by the phase, whose parts are phase[2*axis], phase[2*axis+1] rounded
 * to float32; boundaries.cpp:359-366): a curl or an off-diagonal row reading a neighbour's SAME
 * stored index across the window edge (plane 0) reads the connected word -- conj(phase) below
 * the window, phase above -- as (a*c) - (b*d), (a*d) + (b*c) with each operation rounded on
 * its own (step.cpp:171-221; stepping._bloch_ghost_operand :2545-2586). Linked parts step only
 * together, in one batch (fdtd_batch_create puts both in it). Refused with flux or DFT
 * regions. */
int fdtd_link_parts(fdtd_sim *re, fdtd_sim *im, const int32_t kinds[3], const float phase[6]);

/* A conductivity: sigma (the reference's layout, float32, one volume) for component `axis` of
 * D (field 0, MEEP's D_conductivity) or B (field 1, B_conductivity); null or all zero removes it.
 * Stores condfac = 1 - sigma*dt2 (float32, dt2 = float32(dt)/2) and condinv = 1/(1 + sigma*dt/2)
 * (double, rounded once) and, under an absorber, an f_cond history ("f_cond_Dx" .. "f_cond_Bz"
 * for fdtd_upload/fdtd_download). The curl of that component then takes MEEP's conductive forms.
 * Replaces Fields.set_d_conductivity / set_b_conductivity (reference fields.py:775-870) and
 * stepping._apply_conductive_update / _apply_conductive_pml_update (stepping.py:2820-2983).
 * A non-finite sigma is refused. */
int fdtd_set_conductivity(fdtd_sim *sim, int32_t field, int32_t axis, const float *sigma);

/* The subnormal policy this library was built with (a build-time flag, the reference's
 * floating-point contract, docs/design/floating-point.md): 0 keep (libmeep_fdtd.so), 1 flush
 * (libmeep_fdtd_flush.so: float32 and float64 subnormal inputs read as zero and subnormal
 * results written as zero, sign kept -- x86's FTZ|DAZ, under which MEEP's flush identity set
 * is minted, meep_identity.py:43-52). No GPU call. */
int fdtd_subnormal_mode(void);

/* Host arithmetic under the flush policy: on != 0 sets the calling thread's MXCSR FTZ|DAZ
 * (x86), 0 clears it. For host-side reductions that must round as MEEP's host does under its
 * flush policy (the whole-cell get_array of a flush run, reference readback.py under
 * meep_identity.host_subnormal_mode). No GPU call. */
int fdtd_host_flush(int32_t on);

/* Confine a conductivity set above to the cells whose structure chunk keeps it: cells[q] 0
 * (x fastest, nx*ny*nz bytes) steps the lossless forms there and deposits without the condinv
 * factor, as MEEP does in a chunk whose conductivity samples are all zero (structure.cpp:
 * 883-900). Replaces ChunkLayout.conductive_mask (reference chunk_map.py:720-755) as the
 * stepping reads it. Call after fdtd_set_conductivity for the component. */
int fdtd_set_conductive_cells(fdtd_sim *sim, int32_t field, int32_t axis, const uint8_t *cells);

/* A DFT field region, MEEP's add_dft_fields (DFTMonitor, reference dft.py:1960; the region's
 * block as FdtdDriver.get_dft returns it): the cells [lo, lo+cnt) of the grid, ncomp (1..4)
 * components (codes 0..5 = Ex, Ey, Ez, Hx, Hy, Hz) sampled each accumulating step at the cell
 * centre (yee_grid = 0: MEEP's yee2cent, as for flux planes) or at each component's own Yee site
 * (yee_grid = 1), E at t and H at t - dt/2, and transformed in float64 like a flux plane. It
 * shares the flux planes' ids and read/write/scale calls (its components in place of E1..H2). */
int fdtd_add_dft_region(fdtd_sim *sim, const int32_t lo[3], const int32_t cnt[3], int32_t ncomp,
                        const int32_t *comps, int32_t yee_grid, int32_t nf, const double *omega,
                        double scale, int32_t decimation, int32_t *region);

/* Copy one plane's accumulated transforms to the host as float64 (re, im) pairs, laid out
 * [component: E1, E2, H1, H2 (tangential, cyclic after the normal), or a DFT region's own
 * components in order][frequency][cell, x fastest over the box]. Flushes buffered samples
 * first. */
int fdtd_flux_read(fdtd_sim *sim, int32_t plane, double *out, size_t count);
/* Replace a plane's transforms with `in` (the layout fdtd_flux_read writes), after the samples
 * still buffered are transformed in: MEEP's load_flux_data, which assigns every element and
 * does not add (FluxMonitor.load_dft_data, reference dft.py:3288-3295, _load_flux_planes
 * dft.py:661-733). The steps that follow accumulate on top. */
int fdtd_flux_write(fdtd_sim *sim, int32_t plane, const double *in, size_t count);
/* Multiply a plane's transforms (every component, E and H sides alike) by re + i*im, after the
 * buffered samples are transformed in: FluxMonitor.scale_dfts (reference dft.py:3274-3286);
 * -1 after fdtd_flux_write is MEEP's load_minus_flux_data (dft.py:3297-3316). */
int fdtd_flux_scale(fdtd_sim *sim, int32_t plane, double re, double im);

/* The flux spectrum of one flux plane, reduced on the device from its float64 transforms
 * (the buffered samples are transformed in first): out[f] = sum over the plane's cells of
 * w * Re(E1 conj(H2) - E2 conj(H1)) for each of its count = nf frequencies, w = weights[cell]
 * (host, per box cell, x fastest: the layout fdtd_flux_read writes) or 1 when weights is NULL.
 * The caller scales by the area element and the plane's weight. Replaces the summation of
 * FluxMonitor._integrate_flux (reference dft.py:3107-3130) over downloaded transforms: nf
 * doubles cross to the host instead of 8*nf*cells. Not for DFT field regions. */
int fdtd_flux_spectrum(fdtd_sim *sim, int32_t plane, const double *weights, double *out,
                       size_t count);

/* The same spectrum with nothing allocated: for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops). No counterpart in the reference. `weights` (or NULL for
 * 1) and `out` (count = nf doubles) are DEVICE pointers the caller owns; the result is written
 * in the simulation's stream order (fdtd_synchronize before reading it on the host). The partial
 * sums live in the library's per-GPU static scratch of 1 MiB, so nf is at most 1024 (larger is
 * refused with FDTD_EINVAL); calls that overlap on one GPU (simulations on different streams)
 * race on that scratch and must be ordered by the caller. Stage 1 writes every partial sum
 * stage 2 reads, so no call depends on what an earlier one left there. */
int fdtd_flux_spectrum_noalloc(fdtd_sim *sim, int32_t plane, const double *weights, double *out,
                               size_t count);

/* Batched form: many simulations advanced together, one fused launch per half-step over a
 * work list of (simulation, tile) entries, and one sampling launch for every flux plane of
 * every member. The batch does not own the simulations; while one exists its members are
 * stepped only through it (their own step/upload/download calls return FDTD_ESTATE), and
 * fdtd_batch_destroy hands their counters back. No counterpart in the reference (it steps
 * one simulation per process, README.md:589-590). */
int fdtd_batch_create(fdtd_sim *const *sims, int32_t count, fdtd_batch **out);
void fdtd_batch_destroy(fdtd_batch *batch);
int fdtd_batch_run(fdtd_batch *batch, int32_t nsteps);

/* One half-step of every member: the H pairs (fdtd_batch_step_h), then the D pairs with the
 * counter advance and the flux sampling (fdtd_batch_step_d) -- the batch's fdtd_step_h and
 * fdtd_step_d, the driver's two fused consults for a run on complex storage, whose two parts
 * step only together. */
int fdtd_batch_step_h(fdtd_batch *batch);
int fdtd_batch_step_d(fdtd_batch *batch);
/* Run the float64 transform pass over every plane's buffered samples (before a read). */
int fdtd_batch_flush(fdtd_batch *batch);
int fdtd_batch_synchronize(fdtd_batch *batch);

/* Launch shape (the autotuner sets it; process-wide): vec = x-adjacent cells per thread (1, 2
 * or 4); blocks_per_cu = 0 for one workgroup per tile, else a persistent grid of that many
 * workgroups per CU; xcd_remap = 0 for the plain order, 1 to give each XCD a contiguous eighth of
 * the tile list, sb >= 2 to deal super-blocks of sb tiles round-robin over the 8 XCDs;
 * nt_stores = 1 to stream the fused pairs' outputs with non-temporal stores. No counterpart
 * in the reference (its launches are one thread per cell). */
int fdtd_set_launch(int32_t vec, int32_t blocks_per_cu, int32_t xcd_remap, int32_t nt_stores);
int fdtd_get_launch(int32_t out[4]);  /* the H pair's shape */

/* The same per fused pair: pair 0 = the H pair, 1 = the D pair (each has its own best shape:
 * the D pair streams more arrays and prefers fewer resident workgroups). */
int fdtd_set_launch_pair(int32_t pair, int32_t vec, int32_t blocks_per_cu, int32_t xcd_remap,
                         int32_t nt_stores);
int fdtd_get_launch_pair(int32_t pair, int32_t out[4]);

/* Cells per thread (1, 2 or 4; default 1) of a pair's shell kernel: the rows on a y or z
 * layer line, stepped with every layer form where no simulation runs the general forms (the box
 * of the other rows takes only the x line's forms, with the pair's own width above). The
 * shell's forms hold more operands per cell, so a narrower run keeps more waves resident. */
int fdtd_set_launch_shell(int32_t pair, int32_t vec);
int fdtd_get_launch_shell(int32_t pair, int32_t *vec);

/* The D pair's box rows over LDS-staged halo tiles of `rows` rows (2, 4 or 8; a tile is
 * 256 / rows cells along x, one cell a thread, with the row below and the column left of it
 * staged in LDS), or 0: runs of vec cells; used where the D pair's vec is 1. Process default,
 * as fdtd_set_launch_pair. No counterpart in the reference. */
int fdtd_set_launch_tile(int32_t rows);
int fdtd_get_launch_tile(int32_t *rows);

/* enable = 1 (the default): fdtd_run and fdtd_batch_run replay a captured hipGraph of 128 whole
 * steps (both fused pairs, flux sampling and the transform passes that fall due, each launch
 * reading the step counters at its step's offset, and one counter advance at the end) wherever
 * every flux plane samples each step; 0 launches every step eagerly. Results are
 * the same words either way. Process-wide. No counterpart in the reference (its host loop
 * launches each sub-step). */
int fdtd_set_graphs(int32_t enable);

/* A simulation's (or batch's) own launch configuration, in place of the process defaults the
 * setters above change: cfg[0..5] the H pair's shape (vec, blocks_per_cu, xcd_remap, nt,
 * shell vec, halo tile rows), cfg[6..11] the D pair's, cfg[12] graphs (0/1), cfg[13] the flux
 * transform (0, 1, 2 or 4, as fdtd_set_flux_transform). Halo tile rows: 0, or 2, 4 or 8 for
 * the D pair's box rows over LDS-staged halo tiles of that many rows (vec 1 only). NULL:
 * follow the process defaults again. The get calls write the configuration in force. Results
 * are the same words under every configuration. No counterpart in the reference. */
int fdtd_sim_set_config(fdtd_sim *sim, const int32_t *cfg);
int fdtd_sim_get_config(fdtd_sim *sim, int32_t out[14]);
int fdtd_batch_set_config(fdtd_batch *batch, const int32_t *cfg);
int fdtd_batch_get_config(fdtd_batch *batch, int32_t out[14]);

/* Where some plane decimates, the captured period is the least common multiple of the planes'
 * decimations, repeated up to 128 steps (at most 1024 steps, with fewer than 127 samples of any
 * plane in it): both pairs and the sampling launches, the transform passes run between replays
 * before a ring would overflow; it replays while each plane's decimation phase is the one it
 * was captured at. out[0] = graphs captured, out[1] = replays after their capture, over the
 * simulation's (or batch's) life. No counterpart in the reference. */
int fdtd_graph_stats(fdtd_sim *sim, int64_t out[2]);
int fdtd_batch_graph_stats(fdtd_batch *batch, int64_t out[2]);

/* mfma = 1 (the default), 2 or 4: the float64 flux transform runs as 16x16x4 float64 MFMA
 * products (phase words times ring samples), each wave taking that many 16-cell tiles; 0: the
 * vector-FMA kernel. The sums associate differently (both are float64; they agree to rounding).
 * Process-wide. No counterpart in the reference. */
int fdtd_set_flux_transform(int32_t mfma);

/* Measure average kernel durations over `repeats` launches of each fused pair on the
 * simulation's stream with HIP events: out_ms[0] = H pair, out_ms[1] = D pair. */
int fdtd_time_pairs(fdtd_sim *sim, int32_t repeats, float out_ms[2]);

/* nsteps whole steps launched eagerly in chunks of 128, as fdtd_run launches the steps it does
 * not replay from a graph (each launch reads the step counters at its step's offset in the
 * chunk; the counters advance once per chunk), with HIP events around every fused-pair launch
 * on the simulation's stream, read after each chunk, then the transform pass over every
 * sample still buffered (fdtd_flush), inside the window. out_ms[0] = total elapsed, [1] = mean
 * H pair, [2] = mean D pair (per step). */
int fdtd_run_timed(fdtd_sim *sim, int32_t nsteps, float out_ms[3]);

/* Diagnostic: mean duration of a device copy of `bytes` (the reference point that scales the
 * HBM byte counters of a profile): mode 0 one float4 per thread per pass, 1 four loads in
 * flight per thread, 2 the same with non-temporal stores, 3 read only, 4 write only (the time
 * is then for `bytes` moved once, not twice); one-pass copies with every load issued before the
 * first store: 5 eight float4 loads per thread, non-temporal loads and stores, 6 the same with
 * plain loads and stores, 7 four per thread non-temporal, 8 four per thread with plain loads
 * and non-temporal stores (bytes rounded down to whole 32 KiB groups). No counterpart in the
 * reference. */
int fdtd_copy_probe(uint64_t bytes, int32_t repeats, int32_t mode, float *ms);

/* The fused pairs' access pattern without their arithmetic: `nread` read streams and `nwrite`
 * write streams (<= 16 in all) of n floats, each touched once per thread at the same index,
 * stream q based at q * (n*4 rounded to 2 MiB) + q * `stagger` bytes in one allocation. Mean
 * ms over `repeats` launches. A measurement aid with no counterpart in the reference. */
int fdtd_stream_probe(int32_t nread, int32_t nwrite, int32_t n, uint64_t stagger,
                      int32_t repeats, float *ms);

#ifdef __cplusplus
}
#endif
#endif
