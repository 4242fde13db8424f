/*
 * cgfdm.h — C ABI of the MI355X-native CGFDM time loop.
 *
 * A drop-in for the time loop that the reference drives from run()/propagate():
 * one opaque domain handle owns every device buffer of one simulation (wavefield,
 * packed metric/medium, PML auxiliaries, stations, PGV) on one GPU and one HIP stream.
 * Host buffers belong to the caller; device state belongs to the handle.
 *
 * Conventions
 *  - Every entry point returns an int status (CGFDM_OK == 0, negative on error) and never
 *    throws; cgfdm_last_error() returns a thread-local description of the last failure.
 *  - Grid indices are the reference's global interior indices: X in [0,NX), Y in [0,NY),
 *    Z in [0,NZ) with Z = NZ-1 the free surface (reference station.json / params.json).
 *  - Padded arrays have HALO = 3 cells on each side: _nx_ = NX+6 etc., x fastest
 *    (reference XFAST layout, axis.h:64-66).
 *  - Calls on one handle must be serialised by the caller; different handles are independent.
 */
#ifndef CGFDM_H
#define CGFDM_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFDM_HALO 3

enum cgfdm_status {
    CGFDM_OK = 0,
    CGFDM_EINVAL = -1,   /* bad argument / inconsistent descriptor            */
    CGFDM_ENOMEM = -2,   /* device or host allocation failed (reference 10001) */
    CGFDM_EHIP = -3,     /* HIP runtime error                                  */
    CGFDM_ECFL = -4,     /* DT above the CFL limit (reference calc_CFL, 110)   */
    CGFDM_EPML = -5,     /* nPML larger than the grid (reference codes 110/130)*/
    CGFDM_ESTATE = -6,   /* call out of order (e.g. step before set_model)     */
    CGFDM_EIO = -7,      /* file could not be read or written                  */
    CGFDM_ENONFINITE = -8 /* the wavefield holds a value that is not finite    */
};

/* Storage precision of the wavefield, RK accumulator and packed metric/medium.
 * Arithmetic is always FP32. FP16/BF16 use the reference's nondimensional scaling
 * (macro.h:65-73): V' = Cv*V, T' = Cs*T, buoyancy' = 1/(rho*Cs/Cv*Crho). */
enum cgfdm_precision {
    CGFDM_FP32 = 1,
    CGFDM_FP16 = 2,
    CGFDM_BF16 = 3
};

/* Domain descriptor: the subset of the reference PARAMS/GRID that the time loop reads
 * (struct.h:298-470, getParams.cpp:12-252). */
typedef struct cgfdm_desc {
    int NX, NY, NZ;        /* interior grid points                                   */
    int nPML;              /* PML thickness in cells (reference nPML, default 12)    */
    double DH;             /* grid spacing (m)                                       */
    double DT;             /* time step (s)                                          */
    int NT;                /* number of time steps = int(TMAX/DT) (propagate.cpp:36)  */
    int precision;         /* enum cgfdm_precision                                   */
    int free_surface;      /* 1: traction-free top, no top PML (reference FREE_SURFACE) */
    int pml;               /* 1: auxiliary-variable PML on the boundary faces (PML)  */
    int use_terrain;       /* 1: terrain-following vertical stretch from a height map */
    double depth;          /* model depth (m) used by the stretch (params Depth(km)*1e3) */
    int centerX, centerY;  /* coordinate origin in grid indices (params centerX/Y)   */
    double Cv, Cs, Crho;   /* scale factors; 0 selects the defaults of the precision */
    int check_cfl;         /* 1: reject DT > 1.34*Hmin/Vpmax (calc_CFL.cpp:331-352)  */
    int chi_mode;          /* free-surface matrices: 0 the reference's chi = 2*lambda + mu
                              (contravariant.cpp:266), 1 chi = lambda + 2*mu, the traction-free
                              condition; the two agree where lambda = mu (Vp/Vs = sqrt 3)      */
} cgfdm_desc;

typedef struct cgfdm_domain cgfdm_domain;

/* Returns a description of the last failure on this thread ("" if none). */
const char* cgfdm_last_error(void);

/* Creates a domain and allocates its device state (zeroed wavefield).
 * Replaces allocWave/allocPML/allocPMLParameter/allocCJM/allocMat_rDZ
 * (wave_deriv.cpp:24-48, pml_deriv.cpp:41-72, init_pml_para.cpp:21-51, CJM.cpp:22-39,
 * contravariant.cpp:179-200) and the grid setup of init_grid.cpp:14-130 for one rank. */
int cgfdm_domain_create(const cgfdm_desc* desc, int device, cgfdm_domain** out);

/* Frees every device buffer of the domain (freeWave/freePML/freeCJM...). */
int cgfdm_domain_destroy(cgfdm_domain* h);

/* Device bytes the domain holds (for admission by free HBM). */
long long cgfdm_domain_bytes(const cgfdm_domain* h);

/* Bytes cgfdm_domain_create would allocate for this descriptor (no allocation). */
long long cgfdm_estimate_bytes(const cgfdm_desc* desc);

/* Scale factors of 16-bit storage chosen from the source, for cgfdm_desc.Cv/Cs (Crho unchanged),
 * so that a strong source stays inside the FP16 range where the reference's fixed scales
 * (macro.h:65-73) overflow. Cv/Cs stays at the reference's 5e9, so the stored medium is unchanged.
 *  rate == NULL: the point source (fixed 1e16 N*m Ricker, singleSource.cpp:29-56): the reference's
 *      (5e3, 1e-6) times (DH/100 m)^2.
 *  rate = a moment-rate table as cgfdm_set_source_table takes it (nt x npts x 6, tensor * area *
 *      slip rate, at the global interior indices xyz[3*npts]) sampled every dt s, mu_max the
 *      largest shear modulus (Pa) of the medium: the peak velocity is estimated as the largest,
 *      over nodes p and radii R = DH .. 16 DH, of 2 * sum over the nodes within R of
 *      max_t|rate| / (4 pi R^2) (the field of a patch of nodes slipping together, which tends
 *      to half the slip rate of a planar fault), at R = DH of (2.5 max_t|rate_p| + the sum over
 *      p's neighbours) / (4 pi DH^2) (the near field of one node's moment rate, with the
 *      discrete scheme's excess over the continuum for the node itself), the peak stress as
 *      mu_max * max_t|sum rate dt| / (4 pi DH^3); Cv = min(4e3 / velocity, 5e9 * 1e4 / stress).
 *      The velocity target leaves room for the RK stage values, which overshoot the stored
 *      peak by up to ~4.5x next to a slow single-node source.
 * No counterpart in the reference, whose scales are compile-time constants. */
int cgfdm_source_scales(double DH, const int* xyz, const float* rate, long long npts, int nt, float dt,
                        double mu_max, double* Cv, double* Cs);

/* The same for the explosive point source (singleSource.cpp:29-127: Ricker moment rate of peak
 * frequency fc, M0 = 1e16 N m, over 7^3 Gaussian weights): the rule above applied to the
 * equivalent table, one node per stencil point with rate = s(t) * weight / mu_src on xx, yy, zz.
 * mu_src: shear modulus at the source (or the smallest of the medium), mu_max the largest.
 * No counterpart in the reference. */
int cgfdm_point_source_scales(double DH, double fc, double mu_src, double mu_max, double* Cv, double* Cs);

/* Uploads the model and builds the packed metric/Jacobian/medium array in FP64 on the GPU,
 * the free-surface rDZ matrices and the PML profiles.
 *  terrain : height (m) of the surface on the padded (NY+6)x(NX+6) grid, x fastest;
 *            read only when desc->use_terrain (reference coord.cpp:255-307, terrain.cpp:930-937).
 *  vs,vp,rho: medium on the interior grid, NZ x NY x NX, x fastest (m/s, m/s, kg/m^3).
 * Replaces constructCoord/constructCJM/init_pml_parameter (coord.cpp:312-402,
 * CJM.cpp:170-236, contravariant.cpp:71-467, init_pml_para.cpp:99-285). */
int cgfdm_set_model(cgfdm_domain* h, const double* terrain,
                    const float* vs, const float* vp, const float* rho);

/* Finite-fault source file of the reference (useMultiSource, init_MultiSource,
 * multiSource.cpp:824-922): int32 npts, int32 nt, float32 dt, then per point float32 lon, lat
 * (degrees), z (m), area (m^2), strike, dip (degrees), rake[nt] (degrees), rate[nt] (slip rate).
 * Maps it onto the grid of `desc` (and `terrain`, padded heights, when desc->use_terrain) as the
 * reference does: azimuthal equidistant projection on WGS84 centred on (lon0, lat0) = the
 * grid's (centerX, centerY) (projTrans, :109-173), depths relative to the shallowest point and
 * the grid node of each point's cell (srcCoord2Index, :177-320), moment-rate tensors
 * (solveMomentRate, :478-545), points sharing a node summed (dealDuplicateIndex, :551-607), nodes
 * in increasing order (changeStorageOrder, :752-796). Fills xyz[3*n] (global interior indices)
 * and rate[nt][n][6], the inputs of cgfdm_set_source_table, for the n nodes inside the grid
 * (n <= cap); with xyz == rate == NULL only reports n, nt and dt. Host only. */
int cgfdm_finite_fault_map(const char* path, const cgfdm_desc* desc, const double* terrain, double lon0,
                           double lat0, int* xyz, float* rate, long long cap, long long* npts_out, int* nt_out,
                           float* dt_out);

/* The projection cgfdm_finite_fault_map uses: azimuthal equidistant on WGS84 centred on
 * (lon0, lat0), degrees in, metres out (x east, y north); PROJ's "+proj=aeqd +ellps=WGS84" in
 * the reference (multiSource.cpp:149-170). No counterpart call in the reference's own code. */
int cgfdm_aeqd_forward(double lon0, double lat0, long long n, const double* lon, const double* lat, double* x,
                       double* y);

/* Single explosive Ricker source at a global interior index with peak frequency fc
 * (reference useSingleSource: locateSource/loadPointSource, singleSource.cpp:1-177,
 * propagate.cpp:145-146). M0 = 1e16 N*m as in the reference. */
int cgfdm_set_source_point(cgfdm_domain* h, int X, int Y, int Z, float fc);

/* Multi-point moment-rate source (reference useMultiSource: init_MultiSource,
 * calculateMomentRate, addMomenteRate; multiSource.cpp:824-922, addMoment.cpp:11-339).
 *  xyz : npts x 3 global interior indices (duplicates allowed; they are summed).
 *  rate: nt x npts x 6 moment-tensor rates (xx,yy,zz,xy,xz,yz) * area * slip rate,
 *        i.e. the table after solveMomentRate/changeStorageOrder, before the
 *        mu*Cv/(J*DH^3) scaling, which this call applies on the device. */
int cgfdm_set_source_table(cgfdm_domain* h, const int* xyz, const float* rate,
                           long long npts, int nt, float dt);

/* Stations (reference readStationIndex/initStationIndex, station.cpp:83-296):
 * xyz = n x 3 global interior indices. Points outside the grid are rejected. */
int cgfdm_set_stations(cgfdm_domain* h, const int* xyz, int n);

/* Advances nsteps full time steps (source, 4 RK stages, station and PGV sampling);
 * the body of the loop at propagate.cpp:142-198. Asynchronous on the domain's stream. */
int cgfdm_step(cgfdm_domain* h, int nsteps);

/* Waits for all work queued on the domain. Returns CGFDM_ENONFINITE once the non-finite guard
 * (cgfdm_set_check_interval) has found a wavefield value that is not finite. */
int cgfdm_sync(cgfdm_domain* h);

/* Non-finite guard. No counterpart in the reference, which checks nothing (an FP16 overflow to
 * Inf propagates silently through its time loop, propagate.cpp:142-198). Every `every` steps
 * (default 100; 0 turns it off) cgfdm_step scans the wavefield after the step on the device
 * (one read of the 9 stored planes, ~0.06 % of the step time at the default) and keeps the first
 * step count after which a value was infinite or NaN; cgfdm_sync, cgfdm_get_stations and
 * cgfdm_get_pgv then return CGFDM_ENONFINITE (the outputs are still copied).
 * cgfdm_nonfinite_step waits for the domain and gives that step count, or -1. */
int cgfdm_set_check_interval(cgfdm_domain* h, int every);
int cgfdm_nonfinite_step(cgfdm_domain* h, int* step);

/* Current time-step counter (number of completed steps). */
int cgfdm_current_step(const cgfdm_domain* h);

/* Station seismograms, n x NT x 9 float (Vx,Vy,Vz,Txx,Tyy,Tzz,Txy,Txz,Tyz), physical
 * units; the layout of station_mpi_X_Y_Z.bin after its int32 header (station.cpp:382-405).
 * This and cgfdm_get_pgv return CGFDM_ENONFINITE as cgfdm_sync does. */
int cgfdm_get_stations(cgfdm_domain* h, float* out);

/* Peak ground velocity on the free surface, NY x NX x 2 float (PGVh, PGV);
 * the layout of PGV_Z_mpi_X_Y_Z.bin (PGV.cpp:103-111). */
int cgfdm_get_pgv(cgfdm_domain* h, float* out);

/* Plane of one wavefield variable (0..8), physical units (data_io.cpp:86-231):
 *  axis 0: X = index, out NZ x NY;  axis 1: Y = index, out NZ x NX;
 *  axis 2: Z = index, out NY x NX. */
int cgfdm_get_slice(cgfdm_domain* h, int axis, int index, int var, float* out);

/* cgfdm_get_slice without any device allocation: the plane is staged in a fixed 16 MiB device
 * array of the library (one per GPU, kept until the process ends) instead of a buffer allocated
 * per call, so callers that cannot allocate (tight snapshot loops, code under HIP graph capture,
 * harnesses that forbid hipMalloc) can take snapshots. Planes above 4 Mi points (4194304) are
 * rejected with CGFDM_EINVAL. Calls in one process are serialised on a lock (the staging array is
 * shared by every domain on a GPU and holds the last call's plane); the call waits for the
 * domain's stream. No counterpart in the reference. Same output as cgfdm_get_slice. */
int cgfdm_get_slice_noalloc(cgfdm_domain* h, int axis, int index, int var, float* out);

/* Raw state access for tests and restarts: the 9 wavefield variables on the interior
 * grid (9 x NZ x NY x NX float, x fastest) in storage units (scaled by Cv/Cs). */
int cgfdm_get_state(cgfdm_domain* h, float* out);
int cgfdm_set_state(cgfdm_domain* h, const float* in);

/* The packed metric/Jacobian/medium array on the padded grid, 13 x num float
 * (xi_x..zt_z, J, mu, lambda, buoyancy), decoded from storage. */
int cgfdm_get_cjm(cgfdm_domain* h, float* out);

/* Average duration (ms) of the fused stage kernel launches timed with HIP events on
 * the domain's stream since the last reset, and the number of launches timed. */
int cgfdm_stage_timing(cgfdm_domain* h, int reset, double* avg_ms, long long* count);

/* Per-kernel split of the same timing: out[0] whole stage, out[1] bulk kernel, out[2] boundary
 * (PML and free-surface) kernel, average ms per stage; count = stages timed. */
int cgfdm_kernel_timing(cgfdm_domain* h, int reset, double* out, long long* count);

/* Enables/disables HIP-event timing of the stage kernels (off by default). */
int cgfdm_enable_timing(cgfdm_domain* h, int on);

/* How a stage is split: out[0] points of the bulk box, out[1] the other interior points (PML,
 * free-surface and general boxes),
 * out[2] bulk tile shape (0: 64x8, 1: 64x4, 2: 64x16, 3: 128x8, -1: not tuned yet), out[3] z-chunk
 * (0: automatic). */
int cgfdm_partition(const cgfdm_domain* h, long long* out);

/* Whether the stage kernels take the first metric pair (xi_x, eta_y) from one stored copy:
 * *out = 1 when set_model found the same stored pair at every interior point (the case on the
 * terrain-following grids this library builds, where x depends on i only and y on j only), 0 when
 * the pair is read per point. Results are the same either way. Environment CGFDM_STREAM_PAIR0=1
 * (read at domain creation) forces the per-point reads. A diagnostic of this implementation, with
 * no counterpart in the reference. Waits for the domain's stream. */
int cgfdm_uniform_metric_pair(const cgfdm_domain* h, int* out);

/* The decomposition of the interior a stage uses, from the descriptor alone (no device):
 * writes up to max_boxes records of 8 ints (i0, i1, j0, j1, k0, k1 in padded indices, kind,
 * axis) and returns the number of boxes (or a negative status). kind 0: the bulk box (tile
 * kernel), 1: a single-slab PML box whose normal is `axis` (tile kernel with the auxiliary
 * update), 2: a general box (point kernel: the free-surface layers of the PML columns), 3: the
 * free-surface layers above the bulk columns (free-surface tile kernel), 4: a PML box where two
 * or three slabs meet (an edge or a corner; `axis` is the mask of their normals, bit 0 x, 1 y,
 * 2 z; tile kernel with every slab's auxiliary update). The boxes are disjoint and cover the
 * interior. No counterpart in the reference, which loops over
 * the whole grid in every kernel. */
int cgfdm_decompose(const cgfdm_desc* desc, int* boxes, int max_boxes);

/* Layered 1-D medium following the terrain, built on the device: the rule of the reference's
 * layer-file model (medium.cpp:84-216, LayeredStructureTerrain) for nlayer rows of top depth
 * (m, increasing), Vs, Vp (m/s) and rho (kg/m^3); points above the first interface and
 * below none keep the homogeneous default (3464, 6000, 2670). Otherwise as cgfdm_set_model. */
int cgfdm_set_model_layered(cgfdm_domain* h, const double* terrain, int nlayer, const float* depth,
                            const float* vs, const float* vp, const float* rho);

/* ------------------------------------------------------------------------------------------
 * Serving: many independent simulations on one GPU with continuous batching.
 * No counterpart in the reference, which runs one simulation per MPI job; this is the engine
 * for Green's-function-style request streams. A server owns one stream, one device arena of
 * hbm_budget bytes and a pinned host arena. Requests are admitted in arrival order whenever a
 * slot is free and the arena holds the request's whole device footprint; resident domains are
 * advanced together, steps_per_launch whole steps per launch of one persistent multi-domain
 * kernel set (replayed from a HIP graph when use_graph), and retired when they reach NT.
 * Every domain computes the same numbers alone and in any mix (the per-point arithmetic does
 * not depend on the batch).
 * ------------------------------------------------------------------------------------------ */
typedef struct cgfdm_server cgfdm_server;

typedef struct cgfdm_server_config {
    int device;
    int precision;          /* storage precision of every request                       */
    long long hbm_budget;   /* device arena bytes; 0: 80 % of the free HBM              */
    long long host_budget;  /* pinned host arena bytes; 0: 1 GiB                          */
    int max_domains;        /* resident simulations, 1..64                                */
    int steps_per_launch;   /* whole time steps per batch launch, 1..64                   */
    int use_graph;          /* 1: replay a captured HIP graph; 0: launch directly         */
    int admit_order;        /* which waiting request a free slot takes: 0 the earliest
                               submitted (it waits for room in the arena, holding the
                               others back); 1 the least work (NX*NY*NZ*NT) among those that
                               fit the arena now, except that requests waiting longer than
                               60 s go first, oldest first — a lower median latency when more
                               requests wait than slots, and a bound on how long a large
                               request is passed over                                       */
    int rescale_retries;    /* 16-bit storage: how many times a request whose seismograms or
                               PGV come back non-finite is run again from its start with Cv
                               and Cs divided by 8 (0..8; 0: report CGFDM_ENONFINITE at once) */
} cgfdm_server_config;

typedef struct cgfdm_request {
    cgfdm_desc desc;            /* grid, DH, DT, NT, nPML, ...; precision = the server's;
                                   check_cfl must be 0. With 16-bit storage and Cv = Cs = 0
                                   the server chooses the request's scales from its source
                                   (cgfdm_source_scales, with the medium's largest mu)        */
    const double* terrain;      /* (NY+6) x (NX+6) heights when desc.use_terrain            */
    int nlayer;                 /* > 0: layered medium (as cgfdm_set_model_layered)         */
    const float* layer_depth;
    const float* layer_vs;
    const float* layer_vp;
    const float* layer_rho;
    const float* vs;            /* nlayer == 0: NZ x NY x NX medium (as cgfdm_set_model)    */
    const float* vp;
    const float* rho;
    int source;                 /* 1: explosive point (src_xyz, fc); 2: moment-rate table   */
    int src_xyz[3];
    float fc;
    const int* tab_xyz;         /* as cgfdm_set_source_table                                */
    const float* tab_rate;
    long long tab_npts;
    int tab_nt;
    float tab_dt;
    const int* st_xyz;          /* stations (nst x 3 global indices)                        */
    int nst;
} cgfdm_request;

typedef struct cgfdm_request_info {
    int status;                 /* CGFDM_OK or the failure of this request                 */
    int NT, nst, NX, NY;        /* result shapes: stations nst x NT x 9, PGV NY x NX x 2     */
    double t_submit, t_admit, t_done;  /* seconds since the server started (admit -1: never) */
    long long device_bytes;     /* arena bytes the request held                              */
    char error[160];
    double Cv, Cs;              /* scale factors of the run that produced the results        */
    int rescaled;               /* runs repeated with 8x smaller scales (rescale_retries)    */
} cgfdm_request_info;

typedef struct cgfdm_server_stats {
    long long submitted, completed, failed, launches;
    int resident, pending, peak_resident, graph, tile_cfg, bulk_blocks, general_blocks;
    long long arena_bytes, arena_used, arena_peak;
    long long rescaled;         /* runs repeated with smaller scales after non-finite results */
} cgfdm_server_stats;

/* Creates a server and starts its worker thread. */
int cgfdm_server_create(const cgfdm_server_config* cfg, cgfdm_server** out);

/* Copies the request (all arrays) and queues it; *id identifies it. Fails at once with
 * CGFDM_ENOMEM if the request could never fit the arenas. */
int cgfdm_server_submit(cgfdm_server* s, const cgfdm_request* r, long long* id);

/* Waits up to timeout_ms (< 0: forever) for completed, not yet fetched requests and writes up
 * to max of their ids; returns how many (0 on timeout). */
int cgfdm_server_wait(cgfdm_server* s, long long* ids, int max, int timeout_ms);

/* Copies a completed request's seismograms (nst x NT x 9, as cgfdm_get_stations) and PGV
 * (NY x NX x 2) out (either may be NULL), fills info, releases the record and returns the
 * request's status: CGFDM_ENONFINITE (outputs still copied) when a seismogram or PGV value is not
 * finite after the rescaled runs the configuration allows (rescale_retries; info->Cv, Cs and
 * rescaled describe the run that produced the outputs). */
int cgfdm_server_result(cgfdm_server* s, long long id, float* stations, float* pgv, cgfdm_request_info* info);

/* Counters of the server. */
int cgfdm_server_stats_get(cgfdm_server* s, cgfdm_server_stats* st);

/* Finishes every submitted request, stops the worker and frees everything. */
int cgfdm_server_destroy(cgfdm_server* s);

#ifdef __cplusplus
}
#endif

#endif /* CGFDM_H */
