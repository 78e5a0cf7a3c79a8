/*
 * pn1.h -- C ABI of libpn1, the MI355X (gfx950) force / energy path of
 * Limit-sticky9 (N-body star cluster + central black hole with limited
 * 1PN corrections, Plummer softening, G = 1 units).
 *
 * Conventions shared by every entry point:
 *   - particle 0 is the black hole (BH); n counts it (n = N_stars + 1);
 *   - x, v, a are AoS [n][3] doubles, m is [n] doubles (the reference's
 *     layout, Limit-sticky9.c:742-745);
 *   - c is the speed of light: every 1PN term carries 1/c^2 (the reference
 *     has none, i.e. c = 1, which stays the default of the legacy shapes);
 *   - nothing here takes or returns torch or HIP types.
 *
 * Two precision modes:
 *   PN1_FP64   all pair arithmetic in FP64;
 *   PN1_MIXED  star-star pair arithmetic in FP32 on double-single (hi/lo)
 *              positions with compensated FP32 sums; the BH-star 1PN terms,
 *              the RK update and all energy sums stay in FP64.
 */
#ifndef PN1_H
#define PN1_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (context API) -------------------------------------- */
#define PN1_OK        0
#define PN1_EINVAL   (-1)   /* bad argument or input state                 */
#define PN1_ENOMEM   (-2)   /* device or host allocation failed            */
#define PN1_EDEVICE  (-3)   /* a HIP call failed                           */
#define PN1_ENODEV   (-4)   /* no usable gfx950 device                     */
#define PN1_EIO      (-5)   /* file missing, short or malformed            */

#define PN1_FP64   0
#define PN1_MIXED  1

/* Largest capacity a context or ensemble accepts: the kernels index the
 * six state rows of `capacity` doubles with 32-bit offsets (6 x 2^28 <
 * 2^31); at ~0.5 KB of device state per particle it is also ~130 GB. */
#define PN1_MAX_CAPACITY (1 << 28)

/* ====================================================================== */
/* 1. Legacy shapes: drop-in replacements for the reference's functions.  */
/*    Same prototypes (the reference's [NMAX][3] parameters decay to      */
/*    double (*)[3]); same outputs; energy()/initial_energy() print the   */
/*    same text.  They use a library-owned device context created on     */
/*    first call, the process-global c and mode set below, and on any     */
/*    error print to stderr and exit(1) as the reference does on failure. */
/*    Not thread-safe (neither is the reference).                         */
/* ====================================================================== */

/* Replaces force(), Limit-sticky9.c:21-424.  Overwrites a[0..n); pot is
 * ignored, as in the reference. */
void force(double (*x)[3], double (*v)[3], double *m, double eps,
           double (*a)[3], double *pot, int n);

/* Replaces energy(), Limit-sticky9.c:427-564.  Prints the time, the
 * energy block and DE = (init_ene - E)/E. */
void energy(double t, double (*x)[3], double (*v)[3], double *m, int n,
            double init_ene, double eps);

/* Replaces initial_energy(), Limit-sticky9.c:566-714.  Prints the t = 0
 * block and stores the total in *init_ene. */
void initial_energy(double (*x)[3], double (*v)[3], double *m, int n,
                    double *init_ene, double eps);

/* Speed of light used by the three legacy shapes (default 1.0, which is
 * the reference).  Returns PN1_EINVAL unless c > 0. */
int pn1_set_c(double c);
/* Precision mode of the legacy force() (default PN1_FP64). */
int pn1_set_mode(int mode);

/* ====================================================================== */
/* 2. Context API: device-resident state, no host round trip per stage.   */
/*    Thread-safe across distinct contexts.                               */
/* ====================================================================== */
typedef struct pn1_ctx pn1_ctx;

typedef struct {
    double kin;      /* 1/2 sum m v^2 over all particles incl. the BH      */
    double pot;      /* softened Newtonian potential energy                 */
    double pot_pn;   /* BH-star 1PN energy / c^2                            */
    double pot_pn2;  /* three-body 1PN energy / c^2                         */
    double total;    /* kin + pot + pot_pn + pot_pn2                        */
} pn1_energy_t;

/* Creates a context for up to `capacity` particles on HIP device `device`
 * (1 <= capacity <= PN1_MAX_CAPACITY). */
int  pn1_create(pn1_ctx **out, int capacity, int device);
void pn1_destroy(pn1_ctx *ctx);
/* eps > 0 (Plummer softening), c > 0, dt > 0, mode PN1_FP64 | PN1_MIXED. */
int  pn1_set_params(pn1_ctx *ctx, double eps, double c, double dt, int mode);
/* Copies a state in (1 <= n <= capacity; m[0] > 0; no star on the BH). */
int  pn1_upload(pn1_ctx *ctx, const double *x, const double *v,
                const double *m, int n);
int  pn1_download(pn1_ctx *ctx, double *x, double *v);
/* One force evaluation on the resident state; a is AoS [n][3]. */
int  pn1_force(pn1_ctx *ctx, double *a);
/* The accelerations of stars i0 .. i1 - 1 only, against every particle:
 * the pair kernel runs over those stars' 256-star blocks alone.  This is
 * one rank's work in a run whose stars are split across GPUs (each rank
 * holding every particle's state after an all-gather); the BH, which sums
 * over every star, is not evaluated.  i0 - 1 and i1 - 1 must be multiples
 * of 256 (or i1 = n); a is AoS [i1 - i0][3].  With an explicit j split
 * (pn1_set_tuning, pn1_autotune) it equals rows i0 .. i1 - 1 of pn1_force
 * bit for bit; otherwise the shard gets a split of its own (more chunks
 * for its smaller grid) and agrees to rounding.  No counterpart in the
 * reference, which runs on one processor. */
int  pn1_force_range(pn1_ctx *ctx, int i0, int i1, double *a);
/* Largest n that pn1_force_noalloc takes. */
#define PN1_NOALLOC_MAX_N 16384
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the FP64 force of pn1_force on a state
 * that already lies in device memory.  It allocates nothing: every
 * intermediate lives in a fixed scratch area of the library, 69 MB on each
 * GPU it runs on, held from the library's load there until the process
 * ends.
 *   st      device, 6 rows of n doubles: x, y, z, vx, vy, vz (SoA)
 *   m       device, n masses (m[0] the BH)
 *   acc     device, out: 3 rows of n doubles, ax, ay, az (SoA)
 *   stream  the hipStream_t to enqueue on (NULL: the null stream) of the
 *           current device; the call returns once the kernels are enqueued
 * Requires 1 <= n <= PN1_NOALLOC_MAX_N, finite eps > 0 and c > 0 (else
 * PN1_EINVAL, nothing enqueued); the state itself is not checked.  The j
 * split is 32 chunks, halved while a chunk would hold fewer than 64 stars:
 * pn1_force with that split set (pn1_set_tuning) gives the same bits.  All
 * calls on one GPU share its scratch, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams or threads) race,
 * and the caller serialises them.  The first call on a device looks up the
 * scratch's address; make it before capturing a graph. */
int  pn1_force_noalloc(const double *st, const double *m, int n, double eps,
                       double c, double *acc, void *stream);
/* Energy of the resident state. */
int  pn1_energy(pn1_ctx *ctx, pn1_energy_t *out);
/* `steps` steps of the reference's three-evaluation stage scheme
 * (Limit-sticky9.c:841-893), state kept on the device; the step is
 * captured into a hipGraph on the first call and replayed after that,
 * until pn1_upload, pn1_set_params or pn1_set_tuning changes what it
 * holds (the next call captures it again). */
int  pn1_step(pn1_ctx *ctx, int steps);
/* Blocks until the context's stream is idle. */
int  pn1_sync(pn1_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *pn1_last_error(const pn1_ctx *ctx);

/* Launch shape of the pair kernels: the j range is split into `jsplit`
 * chunks (0 = automatic; a split of 32 or more, a multiple of 8, ends in 16
 * chunks an eighth as long as the others, which the launch runs last, so
 * that it does not wait long on its last workgroups), the mixed kernel
 * compensates its FP32 sums every
 * `tile` pairs (0 = default), and `variant` picks the kernel form (-1 =
 * automatic per mode): bit 0 the j loop (clear: one j-record at a time,
 * set: two in flight), bit 1 (mixed form) one Newton step on the FP32
 * rsqrt, bit 2 (mixed form, splits into a multiple of 4 chunks) four
 * chunks per 1024-thread workgroup combined in LDS instead of one chunk
 * per 256-thread workgroup: a quarter of the partial sums written, ~1.6%
 * more pair-kernel time.
 * pn1_get_tuning reports the values in use.  Used by the tuner. */
int  pn1_set_tuning(pn1_ctx *ctx, int jsplit, int tile, int variant);
int  pn1_get_tuning(const pn1_ctx *ctx, int *jsplit, int *tile, int *variant);
/* Autotuner: times `reps` whole force evaluations of the resident state
 * (every kernel, so the cost of reducing the split's partial sums counts)
 * for each j split (automatic, 8 ... 256) and j loop variant, keeps the
 * fastest (as pn1_set_tuning would) and reports it with its time per
 * evaluation in ms (any output may be NULL).  The mixed form's compensation interval is
 * not changed: it trades accuracy.  No counterpart in the reference. */
int  pn1_autotune(pn1_ctx *ctx, int reps, int *jsplit, int *variant,
                  double *best_ms);

/* Timing of the pair kernel: while enabled, pn1_step/pn1_force launch
 * without a graph and bracket every force pair-kernel launch with HIP
 * events on the context's stream.  pn1_profile_read returns the number
 * of pair launches and their summed duration (ms) since enabling. */
int  pn1_profile_enable(pn1_ctx *ctx, int on);
int  pn1_profile_read(pn1_ctx *ctx, int *launches, double *total_ms);

/* ---- adjoint (FP64) ---------------------------------------------------- */
/* VJP of one force evaluation at the resident state: for a cotangent u
 * (AoS [n][3], one 3-vector per particle's acceleration) returns
 * gx = sum_p u_p . d a_p / d x (AoS [n][3]), gv likewise for v, and
 * *gm0 = sum_p u_p . d a_p / d m[0] (gm0 may be NULL).  No counterpart in
 * the reference (it has no derivatives); the derivative of force() as
 * written, Limit-sticky9.c:21-424, at this context's c. */
int  pn1_force_vjp(pn1_ctx *ctx, const double *u, double *gx, double *gv,
                   double *gm0);
/* Gradient of a loss of the state after `steps` steps of the stage driver
 * (Limit-sticky9.c:841-893) with respect to the initial state and the BH
 * mass, by the discrete adjoint of the driver.  pn1_adjoint_forward steps
 * the resident state (FP64 kernels whatever the mode) and keeps a
 * checkpoint every `interval` steps; the final state stays resident (read
 * it with pn1_download, evaluate the loss and its gradient on the host).
 * pn1_adjoint_backward takes that gradient (gx, gv: d loss / d final x, v,
 * AoS [n][3]) and returns d loss / d initial x, v (gx0, gv0) and d loss /
 * d m[0] through the dynamics (*gm0, may be NULL).  It recomputes each
 * segment from its checkpoint, keeping the segment's stage states (device
 * memory: ceil(steps / interval) + 3 interval states of 6 x capacity
 * doubles), and leaves the initial state resident.  It fails with
 * PN1_EINVAL if the state or eps, c, dt changed in between. */
int  pn1_adjoint_forward(pn1_ctx *ctx, int steps, int interval);
int  pn1_adjoint_backward(pn1_ctx *ctx, const double *gx, const double *gv,
                          double *gx0, double *gv0, double *gm0);

/* ====================================================================== */
/* 2b. Ensemble API: many independent clusters (each a BH + its stars,    */
/*     as one run of Limit-sticky9.c) advanced together on one device.    */
/*     No counterpart in the reference, which runs one cluster per        */
/*     process; every cluster's physics is the context API's.             */
/* ====================================================================== */
/* Cluster c is particles [offsets[c], offsets[c+1]) of the AoS arrays;
 * its first particle is its BH.  All clusters share eps, c, dt and mode.
 * One force evaluation of the ensemble is one persistent pair-kernel
 * launch: workgroups take (cluster, 256-star block, j chunk) items from
 * eight per-XCD queues (whole clusters per queue, balanced by cost;
 * largest items first; idle workgroups take from the other queues); each
 * item writes its own partial sums, so results do not depend on the order
 * items run in. */
typedef struct pn1_ens pn1_ens;

/* Up to `capacity` particles in total (at most PN1_MAX_CAPACITY) in up to
 * `max_clusters` clusters. */
int  pn1_ens_create(pn1_ens **out, int capacity, int max_clusters,
                    int device);
void pn1_ens_destroy(pn1_ens *ens);
int  pn1_ens_set_params(pn1_ens *ens, double eps, double c, double dt,
                        int mode);
/* offsets[0] = 0, offsets[nclusters] = total particle count; every
 * cluster is checked as pn1_upload checks one state. */
int  pn1_ens_upload(pn1_ens *ens, int nclusters, const int *offsets,
                    const double *x, const double *v, const double *m);
int  pn1_ens_download(pn1_ens *ens, double *x, double *v);
/* Accelerations of every particle, AoS [n_total][3]. */
int  pn1_ens_force(pn1_ens *ens, double *a);
/* Per-cluster energies: out[5 c + k] = {kin, pot, pot_pn, pot_pn2, total}
 * (the fields of pn1_energy_t) of cluster c. */
int  pn1_ens_energy(pn1_ens *ens, double *out);
/* `steps` steps of every cluster (one hipGraph per step, replayed). */
int  pn1_ens_step(pn1_ens *ens, int steps);
int  pn1_ens_sync(pn1_ens *ens);
const char *pn1_ens_last_error(const pn1_ens *ens);
/* Work-item shape: j chunk length `chunk` (records per item; 0 = chosen
 * at upload so that each persistent workgroup takes about 8 items; applies
 * from the next upload), mixed compensation interval `tile` and kernel
 * `variant` as pn1_set_tuning. */
int  pn1_ens_set_tuning(pn1_ens *ens, int chunk, int tile, int variant);
int  pn1_ens_get_tuning(const pn1_ens *ens, int *chunk, int *tile,
                        int *variant);
/* Autotuner of the ensemble, as pn1_autotune, over chunk lengths
 * 256 ... 1024 and the j loop variant; applies at once (the work tables
 * are rebuilt) and invalidates a pending adjoint. */
int  pn1_ens_autotune(pn1_ens *ens, int reps, int *chunk, int *variant,
                      double *best_ms);
/* Adjoint, per cluster, as pn1_force_vjp / pn1_adjoint_forward /
 * pn1_adjoint_backward (FP64 whatever the mode); gm0 (may be NULL) gets
 * one value per cluster: the derivative with respect to that cluster's BH
 * mass.  No counterpart in the reference. */
int  pn1_ens_force_vjp(pn1_ens *ens, const double *u, double *gx,
                       double *gv, double *gm0);
int  pn1_ens_adjoint_forward(pn1_ens *ens, int steps, int interval);
int  pn1_ens_adjoint_backward(pn1_ens *ens, const double *gx,
                              const double *gv, double *gx0, double *gv0,
                              double *gm0);
/* Pair-kernel timing, as pn1_profile_enable / pn1_profile_read. */
int  pn1_ens_profile_enable(pn1_ens *ens, int on);
int  pn1_ens_profile_read(pn1_ens *ens, int *launches, double *total_ms);

/* ====================================================================== */
/* 3. Host-side helpers (no device needed).                               */
/* ====================================================================== */
/* Validates a state as the context API does; writes a reason to msg. */
int pn1_check_input(const double *x, const double *v, const double *m,
                    int n, double eps, char *msg, int msglen);

/* Initial conditions (all positions / velocities FP32-exact, BH at the
 * origin at rest, stars of mass 1/N).  SplitMix64 stream from `seed`.
 *   uniform sphere of radius R, Gaussian velocities scaled to virial
 *   ratio Q = K/|W| (W incl. the BH, softened with eps);
 *   Plummer sphere (scale radius a, total star mass 1), Aarseth-Henon-
 *   Wielen sampling with the mass fraction cut at 0.999.               */
int pn1_ic_uniform_sphere(int n_stars, unsigned long long seed, double R,
                          double Q, double m_bh, double eps,
                          double *x, double *v, double *m);
int pn1_ic_plummer(int n_stars, unsigned long long seed, double a,
                   double m_bh, double *x, double *v, double *m);

/* .dat input of Limit-sticky9.c:795-809: n-1 lines "x y z", then n-1
 * lines "vx vy vz", read as float; the BH row is not in the file and is
 * set to zero.  Masses are not read. */
int pn1_read_dat(const char *path, int n, double *x, double *v);
/* Snapshot of Limit-sticky9.c:820-826: n lines "%lf %lf %lf \n" of
 * positions then n of velocities (BH first).  append != 0 appends. */
int pn1_write_snapshot(const char *path, int append, const double *x,
                       const double *v, int n);
/* Exact binary checkpoint (header + FP64 x, v, m) for resume. */
int pn1_write_checkpoint(const char *path, int n, double t, long long step,
                         double eps, double c, double dt, const double *x,
                         const double *v, const double *m);
/* Reads a checkpoint; pass x = v = m = NULL to read only the header. */
int pn1_read_checkpoint(const char *path, int *n, double *t,
                        long long *step, double *eps, double *c, double *dt,
                        double *x, double *v, double *m, int capacity);

#ifdef __cplusplus
}
#endif
#endif /* PN1_H */
