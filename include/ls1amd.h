/*
 * ls1amd.h -- C ABI of the MI355X short-range force path for ls1-MarDyn.
 *
 * One context owns one HIP stream on one device and the device-resident
 * molecule state. Calls on one context must be serialised by the caller (the
 * reference's CellProcessor is driven from one traversal at a time). Every
 * call returns MDX_OK (0) or a negative status; mdx_last_error() gives the
 * message. No exceptions cross the ABI. Arrays are caller-owned host memory,
 * borrowed for the duration of the call; the mdx_*_device entry points take
 * device pointers instead.
 *
 * Component records use the ls1 ASCII phase-space header layout
 * (src/io/ASCIIReader.cpp:152-232), flattened to doubles, per component:
 *   nlj nc nd nq
 *   nlj x (x y z m eps sigma rc_shift do_shift)
 *   nc  x (x y z m q)
 *   nd  x (x y z ex ey ez absMy)
 *   nq  x (x y z ex ey ez absQ)
 *   I11 I22 I33        (values > 0 override the moments computed from sites)
 * Mixing coefficients are (xi, eta) per component pair i<j, in the order
 * (0,1),(0,2),..,(1,2),.. (ASCIIReader.cpp:213-232).
 *
 * Reference interfaces replaced (paths relative to the ls1-MarDyn source):
 *   CellProcessor            src/particleContainer/adapter/CellProcessor.h:29-94
 *   VectorizedCellProcessor  src/particleContainer/adapter/VectorizedCellProcessor.h:29-321
 *   ParticleContainer        src/particleContainer/ParticleContainer.h:85-279
 *   LinkedCells              src/particleContainer/LinkedCells.cpp
 *   Leapfrog                 src/integrators/Leapfrog.cpp:32-156
 */
#ifndef LS1AMD_H
#define LS1AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK 0
#define MDX_ERR_ARG (-1)      /* bad argument / unsupported configuration */
#define MDX_ERR_HIP (-2)      /* HIP runtime error */
#define MDX_ERR_STATE (-3)    /* call out of order (e.g. no model yet) */
#define MDX_ERR_CAPACITY (-4) /* an internal capacity would be exceeded */
#define MDX_ERR_NODEVICE (-5) /* no HIP device */

/* precision forms */
/* boundary types per dimension (BoundaryUtils::BoundaryType,
 * src/parallel/boundaries/BoundaryUtils.h:29) */
#define MDX_BOUNDARY_PERIODIC 0
#define MDX_BOUNDARY_OUTFLOW 1
#define MDX_BOUNDARY_REFLECTING 2
#define MDX_PREC_FP64 0       /* all arithmetic FP64 (reference PRECISION=DOUBLE) */
#define MDX_PREC_FP32_FP64ACC 1 /* FP32 pair arithmetic, FP64 accumulation,
                                   FP64 sums and FP64 Leapfrog (PRECISION=MIXED) */

/* macro[] layout: the four sums VectorizedCellProcessor::endTraversal reduces
 * (VectorizedCellProcessor.cpp:122-155). Upot = macro[0]/6 + macro[1] +
 * macro[3]; Virial = macro[2] + 3 * macro[3]. */
#define MDX_UPOT6LJ 0
#define MDX_UPOTXPOLES 1
#define MDX_VIRIAL 2
#define MDX_MYRF 3

typedef struct mdx_context mdx_context;

/* Thread-local message of the last failed call on ctx (or of mdx_create). */
const char *mdx_last_error(const mdx_context *ctx);

/* Creates a context on HIP device `device` (its own stream, default priority).
 * Replaces the construction of VectorizedCellProcessor + LinkedCells
 * (Simulation.cpp:471-519, 851-862). */
int mdx_create(int device, mdx_context **out);
int mdx_destroy(mdx_context *ctx);

/* Component model and parameter tables: Component::addLJcenter etc.
 * (Component.cpp:103-216), Comp2Param::initialize (Comp2Param.cpp:11-202) and
 * the VCP table build (VectorizedCellProcessor.cpp:19-96). eps_rf feeds
 * epsRFInvrc3 (VectorizedCellProcessor.cpp:21). mix may be NULL (all 1). */
int mdx_set_components(mdx_context *ctx, int ncomp, const double *desc, int desc_len,
                       const double *mix, double eps_rf);

/* Replaces the LJ tables with caller tables (24 eps, sigma^2, 6 * shift) over
 * all LJ centres, row/column = LJ lookup id, as VectorizedCellProcessor's
 * _eps_sig/_shift6 (VectorizedCellProcessor.cpp:56-80). n_ljc must equal the
 * model's LJ centre count. */
int mdx_set_lj_tables(mdx_context *ctx, int n_ljc, const double *eps24, const double *sig2,
                      const double *shift6);

/* Domain [0,box) per dimension, cutoffs and cell geometry.
 * CellProcessor(rc, rcLJ) (CellProcessor.h:37), LinkedCells(bBox, rc)
 * (LinkedCells.cpp:38-106, cell_mode 0, double rc) or LinkedCells::rebuild
 * (LinkedCells.cpp:132-201, cell_mode 1, float rc). periodic != 0: periodic
 * boundaries in all dimensions (DomainDecompBase.cpp:80-376); 0: open
 * boundaries as in the reference's ParticleContainerFactory test set-up. */
int mdx_set_domain(mdx_context *ctx, const double *box, double rc, double rc_lj, int periodic,
                   int cell_mode);

/* Batched replicas: nrep independent systems that share the component model,
 * the cutoffs and the boundary mode, each in its own box [0, boxes[3r..3r+2])
 * with its own linked cells (the LinkedCells rules of mdx_set_domain per
 * replica), stored back to back in one context and stepped together: every
 * kernel of a step is one launch over all replicas (the pair kernel walks the
 * flattened cell table). counts[r] molecules belong to replica r; mdx_upload
 * then takes sum(counts) molecules, replica by replica. Per-replica outputs:
 * macro [nrep][4] (mdx_traverse), sums [nrep][4] (mdx_leapfrog_post),
 * per_step [nsteps][nrep][6] (mdx_step). No counterpart in the reference,
 * which runs one system per process; each replica is what one ls1 run with
 * the same input computes. mdx_cell_ids / mdx_cell_geometry / mdx_cell_pairs
 * / mdx_adjoint serve single systems only. */
int mdx_set_replicas(mdx_context *ctx, int nrep, const double *boxes, const int64_t *counts,
                     double rc, double rc_lj, int periodic, int cell_mode);

/* Launch shape of the pair kernels: cells_per_block consecutive cells per
 * work item (1..64), and blocks_per_xcd persistent 64-lane blocks per XCD
 * walking the items (0: one block per item). Default (1, 0). */
int mdx_set_pair_config(mdx_context *ctx, int cells_per_block, int blocks_per_xcd);

/* Times the pair kernels of the current state over the candidate launch
 * shapes (cells_per_block 1, 2, 4, 8 x blocks_per_xcd 0, 512, 1024),
 * reps timed traversals each (HIP events; a shape's time is its fastest
 * run), and keeps the fastest shape. chosen (may be NULL) =
 * {cells_per_block, blocks_per_xcd}; best_ms = its time. */
int mdx_autotune(mdx_context *ctx, int reps, int32_t *chosen, double *best_ms);

/* The NVT tail of Simulation::simulateOneTimestep (Simulation.cpp:1161-1266)
 * inside every mdx_step, per replica (nrep = the context's replica count):
 *  - lrc != 0: the homogeneous long-range correction (Homogeneous::init /
 *    calculateLongRange, src/longRange/Homogeneous.cpp:21-146): VirialCorr /
 *    (3N) is added to every molecule's virial diagonal before calcFM's site
 *    sums, and UpotCorr / VirialCorr to the step record's Upot and virial
 *    (Domain::calculateGlobalValues, Domain.cpp:187-188). Only unshifted LJ
 *    table entries contribute, as in the reference. The default (0) matches
 *    <longrange type="none">.
 *  - T_target[r] > 0: after upd_postF, the global thermostat's temperature
 *    and scaling factors beta_trans = (3 N Ti / sum m v^2)^0.4, beta_rot =
 *    (rotDOF Ti / sum I w^2)^0.4 (Domain.cpp:219-259, Tfactor 1), the
 *    explosion heuristic when a beta is below 0.9 and explosion_heuristics
 *    != 0 (Domain.cpp:261-313), and VelocityScalingThermostat::apply
 *    (VelocityScalingThermostat.cpp:70-90): v *= beta_trans, D *= beta_rot.
 *    T_target NULL or all <= 0: no thermostat (the NVE path, the default).
 * Settings are reset by mdx_set_domain / mdx_set_replicas. */
int mdx_set_thermostat(mdx_context *ctx, int nrep, const double *T_target, int lrc,
                       int explosion_heuristics);

/* Boundary type per dimension, replacing the <boundaries> section
 * (Simulation.cpp:429-444) after mdx_set_domain / mdx_set_replicas:
 * MDX_BOUNDARY_PERIODIC (images across the box, wrap), or a wall with no
 * images in that dimension (the reference deletes those halos,
 * BoundaryHandler::removeNonPeriodicHalos, BoundaryHandler.cpp:143-170) where,
 * before every upd_preF, BoundaryHandler::processGlobalWallLeavingParticles
 * (BoundaryHandler.cpp:82-134) handles a molecule within rc of the wall that
 * would cross it: MDX_BOUNDARY_REFLECTING reverses its velocity,
 * MDX_BOUNDARY_OUTFLOW deletes it. Outflow walls serve single systems without
 * the thermostat, the long-range correction or the reduced-memory mode (they
 * keep the molecule count fixed); with them mdx_download and mdx_cell_ids
 * return the surviving molecules in upload order (ids tell which), and
 * mdx_molecule_count gives their number. */
int mdx_set_boundaries(mdx_context *ctx, const int32_t *types);

/* Molecules in the context: the uploaded count, less those deleted at outflow
 * walls (ParticleContainer::getNumberOfParticles, ParticleContainer.h:160). */
int mdx_molecule_count(mdx_context *ctx, int64_t *n);

/* Reduced-memory one-centre mode: the VCP1CLJRMM processor with LeapfrogRMM
 * (VCP1CLJRMM.cpp:300-362, LeapfrogRMM.cpp:42-107, PositionCellProcessorRMM.h,
 * VelocityCellProcessorRMM.h). Velocities live at half steps: this call runs
 * the initial force traversal with dt/m halved (Simulation.cpp:902-937), and
 * every following mdx_step (same dt) moves r += dt v, re-bins, and lets the
 * pair kernel add dt/m F to v directly (no F / Vi stored); the step record's
 * sum m v^2 becomes the reference's sum v^2 (no mass factor) of the half-step
 * velocities, sum I w^2 = 0. One component of one LJ centre only (as the
 * reference's reduced-memory build). While the mode is on, mdx_traverse is
 * VCP1CLJRMM's traversal too (force and kick), and mdx_leapfrog_pre /
 * mdx_leapfrog_post are LeapfrogRMM's position / velocity-sum halves.
 * dt = 0 switches back (velocities stay where they are); a new upload also
 * ends the mode. */
int mdx_set_rmm(mdx_context *ctx, double dt);

/* Per-replica record of the last step's NVT tail, out[nrep][8] = {T_target,
 * UpotCorr, VirialCorr, VirialCorr/(3N), T, beta_trans, beta_rot, explosion
 * (0/1)}; T is computed before the velocities are scaled (Domain.cpp:226-230). */
int mdx_thermostat_state(mdx_context *ctx, double *out);

/* MDX_PREC_FP64 (default) or MDX_PREC_FP32_FP64ACC. */
int mdx_set_precision(mdx_context *ctx, int precision);

/* Molecule state in ls1's ICRVQD order: positions r[3n], velocities v[3n],
 * quaternions q[4n] (qw qx qy qz), angular momenta L[3n] (lab frame, the
 * reference's D), component ids cid[n] (0-based). ids may be NULL.
 * Replaces ParticleContainer::addParticles (ParticleContainer.h:121). */
int mdx_upload(mdx_context *ctx, int64_t n, const double *r, const double *v, const double *q,
               const double *L, const int32_t *cid, const uint64_t *ids);

/* Copies the state back in upload order. Any pointer may be NULL. F, M and Vi
 * are the molecule force, torque and virial of the last traversal
 * (FullMolecule::calcFM, FullMolecule.cpp:548-631). q is normalised once per
 * container update since the upload, as setupSoACache does
 * (FullMolecule.cpp:719). */
int mdx_download(mdx_context *ctx, int64_t n, double *r, double *v, double *q, double *L,
                 double *F, double *M, double *Vi, uint64_t *ids);

/* Periodic wrap + counting-sort re-bin + SoA site build.
 * Replaces LinkedCells::update (LinkedCells.cpp:241-354),
 * DomainDecompBase::exchangeMolecules (DomainDecompBase.cpp:80-114) and
 * LinkedCells::updateMoleculeCaches / FullMolecule::setupSoACache
 * (LinkedCells.cpp:1097-1107, FullMolecule.cpp:711-767). */
int mdx_update_container(mdx_context *ctx);

/* Force traversal: every site pair of every molecule pair closer than the
 * cutoff, then FullMolecule::calcFM. A state not binned since its upload is
 * binned first (as are mdx_autotune's, mdx_adjoint's and mdx_cell_ids'). Replaces LinkedCells::traverseCells with
 * VectorizedCellProcessor (LinkedCells.cpp:565-575: initTraversal,
 * processCell/processCellPair over the C08 enumeration, endTraversal) and
 * Simulation::updateForces (Simulation.cpp:835-845). macro may be NULL. */
int mdx_traverse(mdx_context *ctx, double *macro);

/* Leapfrog::transition1to2 -> FullMolecule::upd_preF (Leapfrog.cpp:45-61,
 * FullMolecule.cpp:331-361). */
int mdx_leapfrog_pre(mdx_context *ctx, double dt);

/* Leapfrog::transition2to3 -> FullMolecule::upd_postF (Leapfrog.cpp:63-147,
 * FullMolecule.cpp:363-387). sums (may be NULL) = {sum m v^2, sum I w^2, N,
 * rotational DOF}. */
int mdx_leapfrog_post(mdx_context *ctx, double dt, double *sums);

/* nsteps of Simulation::simulateOneTimestep's force path (Simulation.cpp:
 * 1076-1307 without thermostat / long-range correction / plugins):
 * upd_preF, wrap + re-bin, traversal, calcFM, upd_postF. The context must hold
 * the forces of the current state (call mdx_update_container + mdx_traverse
 * once first, as Simulation::prepare_start does). per_step (may be NULL) gets
 * 6 doubles per step: Upot, virial, sum m v^2, sum I w^2, Upot_LJ, Upot_rest.
 * use_graph != 0 replays one captured hipGraph per step. Without a thermostat
 * (mdx_set_thermostat) step n's upd_postF and step n+1's upd_preF run as one
 * pass over the molecules, with the same results as one-step calls. */
int mdx_step(mdx_context *ctx, int nsteps, double dt, double *per_step, int use_graph);

/* Blocks until the context's stream is idle. */
int mdx_synchronize(mdx_context *ctx);

/* ls1 cell index (LinkedCells::getCellIndexOfPoint, LinkedCells.cpp:870-923,
 * halo layer included) of every molecule, in upload order. */
int mdx_cell_ids(mdx_context *ctx, int64_t n, int64_t *out);

/* Cell geometry: cells per dimension including the halo layer, cell length. */
int mdx_cell_geometry(mdx_context *ctx, int32_t *cells_per_dim, double *cell_length);

/* The cell pairs the traversal evaluates, as ls1 would hand them to
 * processCell / processCellPair (C08CellPairTraversal.h:45-75,
 * C08BasedTraversals.h:46-149), expressed on ls1's halo-extended cell grid:
 * pairs[2k], pairs[2k+1]. Returns the count through *count; writes at most
 * max_pairs. */
int mdx_cell_pairs(mdx_context *ctx, int32_t *pairs, int64_t max_pairs, int64_t *count);

/* Work counters of the last traversal (FlopCounter categories,
 * FlopCounter.cpp:19-76): [0] molecule-centre distance tests (each unordered
 * pair once), [1..7] site pairs inside the cutoff: LJ, C-C, C-D, D-D, C-Q,
 * D-Q, Q-Q. Counting is enabled by mdx_set_counting(ctx, 1). */
int mdx_set_counting(mdx_context *ctx, int enable);
int mdx_counters(mdx_context *ctx, int64_t *out);

/* Per pair kernel (one per i-site type: LJ, charge, dipole, quadrupole), for
 * the roofline of each: ms[4] = its duration in the last traversal timed with
 * mdx_set_timing (HIP events around each launch; 0 for an absent type);
 * counters[4][8] = its raw visit counts in the last counted traversal, in the
 * mdx_counters categories. Every site evaluates its own side of a pair, so an
 * unordered pair of two types is visited once by each type's kernel and a
 * pair of one type twice by its kernel; mdx_counters is half the sum over the
 * kernels. Either pointer may be NULL. No counterpart in the reference (its
 * FlopCounter counts one processor over all types, FlopCounter.cpp:19-76). */
int mdx_kernel_stats(mdx_context *ctx, double *ms, int64_t *counters);

/* Device-resident variants for callers that keep the state in HBM: device
 * pointers to n doubles / ints laid out as in mdx_upload. */
int mdx_upload_device(mdx_context *ctx, int64_t n, const double *r, const double *v,
                      const double *q, const double *L, const int32_t *cid);

/* Time of the last traversal's pair kernels in milliseconds, measured with
 * HIP events on the context's stream (0 if not measured); after an adjoint
 * call with timing on: pair_ms = its derivative kernels, total_ms = up to the
 * assembled VJP (host transfers excluded). */
int mdx_set_timing(mdx_context *ctx, int enable);
int mdx_last_kernel_ms(mdx_context *ctx, double *pair_ms, double *total_ms);

/* The LJ tables in use ([n_ljc][n_ljc] each; any pointer may be NULL):
 * eps24 = 24 eps_ab, sig2 = sigma_ab^2, shift6, as
 * VectorizedCellProcessor's _eps_sig / _shift6 (VectorizedCellProcessor.cpp:
 * 40-80), and shift_rc = the radius at which shift6 is built from the entry's
 * own (eps, sigma) (Comp2Param.cpp:27-38,79-82,102-105), 0 where shift6 is
 * zero or borrowed from another site. */
int mdx_lj_tables(mdx_context *ctx, int n_ljc, double *eps24, double *sig2, double *shift6,
                  double *shift_rc);

/* Charge of every charge site, q[comp * 8 + site] (replaces the charge
 * records that Component::addCharge stores, Component.cpp:129-137, and that
 * FullMolecule::setupSoACache copies into _charges_q, FullMolecule.cpp:
 * 711-767). Site arrays of an uploaded state are rebuilt at once. A
 * parameter-fitting caller uses it between adjoint evaluations. */
int mdx_set_charges(mdx_context *ctx, int ncomp, const double *q);

/* FP64 adjoint of the force path for LJ and charge sites (no counterpart in
 * the reference, whose force path has no derivatives; it differentiates what
 * VectorizedCellProcessor::_loopBodyLJ / _loopBodyCharge compute,
 * VectorizedCellProcessor.cpp:171-277, over the same cell pairs). Needs the
 * state of the last mdx_update_container. All outputs are optional (NULL).
 *   dU_deps, dU_dsig, dW_deps, dW_dsig : [nlj][nlj] derivatives of
 *       U = upot6/6 + upotXpoles and W = virial with respect to the
 *       symmetric pair parameters eps_ab, sigma_ab of LJ lookup ids a, b
 *       (eps24 = 24 eps, sig2 = sigma^2); the truncation shift moves with
 *       them where Comp2Param ties it to them (a site with itself when
 *       shifted, and every cross-component pair at rcLJ);
 *   dU_dq, dW_dq : [ncomp][8] derivatives with respect to each charge;
 *   vjp : [n][3] (upload order) sum_j c_j . dF_j/dr_i for the cotangent
 *       cot [n][3] on the molecule forces (orientations held fixed).
 * Models with dipole or quadrupole sites, or more than 16 LJ centres, are
 * rejected with MDX_ERR_ARG / MDX_ERR_CAPACITY. */
int mdx_adjoint(mdx_context *ctx, int64_t n, const double *cot, double *dU_deps,
                double *dU_dsig, double *dW_deps, double *dW_dsig, double *dU_dq,
                double *dW_dq, double *vjp);

/* mdx_adjoint without any allocation. No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight fitting or serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as mdx_adjoint, but its device buffers
 * (parameter partials, reduced tables, per-site VJP terms, the reordered
 * cotangent) live in a 64 MiB scratch array that the library holds once per
 * device until the process ends, instead of buffers mdx_adjoint allocates on
 * its first call. A state whose buffers do not fit is refused with
 * MDX_ERR_CAPACITY (about 1.3 M one-centre molecules with a cotangent; any
 * size without one). Every call on a device shares that device's array:
 * calls that overlap on one device (several contexts driven from several
 * host threads) race and must be serialised by the caller. */
int mdx_adjoint_noalloc(mdx_context *ctx, int64_t n, const double *cot, double *dU_deps,
                        double *dU_dsig, double *dW_deps, double *dW_dsig, double *dU_dq,
                        double *dW_dq, double *vjp);

#ifdef __cplusplus
}
#endif
#endif
