/* hem_internal.h -- host-side state of the C library (not installed). */
#ifndef HEM_INTERNAL_H_
#define HEM_INTERNAL_H_

#include <pthread.h>
#include <stdint.h>

#include "../../../include/hera_em.h"
#include "../dev/hem_dev.h"

/* Iterations per captured graph of the plain solve and of the batched
 * solve.  Iterations past convergence are no-ops (every kernel reads the
 * control block's done flag first). */
#define HEM_ITERS_PER_GRAPH 8
#define HEM_BATCH_ITERS_PER_GRAPH 4
/* Monotonicity fallback: an iteration keeps u2 when ell(lam) - ell(u2) <
 * -HEM_FB_TOL times the sum of the difference's terms' magnitudes. */
#define HEM_FB_TOL 1e-10
/* Distinct option sets whose graphs a handle keeps. */
#define HEM_GRAPH_CACHE 8
/* device arrays of a tiling (hem_tiles_upload) */
#define HEM_TILE_ARRAYS 10

/* Workspace of a batched solve, kept on the handle and grown on demand. */
typedef struct {
	uint32_t R;                    /* capacity in replicates            */
	double *A, *B, *C;             /* [tile position][R] state          */
	uint32_t *cat_rm;              /* [R][K] counts, caller order       */
	uint32_t *catT;                /* [K][R]: unique rows, class rows   */
	uint32_t *uq;                  /* [tile position][R] unique counts  */
	uint32_t *cnt;                 /* [tile position][R] class counts   */
	double *part;                  /* [tile][R][4] norm partials        */
	double *X;                     /* [2][R][T] r and v rows (lam rows
	                                  at the end)                       */
	uint32_t *m;                   /* [R][B] level-1 resample counts    */
	uint32_t *rs_plan;             /* K weights + chunk starts + prefix */
	uint64_t *seeds;               /* [R]                               */
	uint32_t *u32;                 /* k, done, conv, notconv, active,
	                                  alist, exact-sum reruns           */
	double *f64;                   /* alpha, mean_r, mean_v, lam_sum,
	                                  q_r/q_v, exact pass-1 sums        */
	double *sc;                    /* exact sums: [2][2R][P] chunk sums
	                                  and drifts                        */
	long long *sbase;              /* [2R][P] window bases              */
	double *stab;                  /* [rows][P][W] candidates           */
	uint8_t *snx;                  /* [rows][P][W] window links         */
	double *swa, *swlh;           /* [2R][P][L/2] chunk programs       */
	uint32_t *swrg, *swcnt;
	unsigned long long *scpub;
	double *scsq;   /* pass-1 chunk sums of squares (pass 2's seeds) */
	uint32_t *h_active;            /* pinned                            */
	/* captured iterations; the loop bounds and stop thresholds are kernel
	 * arguments, so a graph is reused only for identical options */
	void *graph[HEM_GRAPH_CACHE];
	int graph_key[HEM_GRAPH_CACHE];
	uint32_t graph_R[HEM_GRAPH_CACHE];
	hem_opts graph_opts[HEM_GRAPH_CACHE];
	int n_graph;
} hem_batch_ws;

/* Workspace of the gradient, HVP and standard-error calls. */
typedef struct {
	int ready;                     /* T-sized arrays allocated          */
	double *theta, *ell, *pi, *w, *x, *E, *fd, *gth, *gell, *a, *v, *hv;
	float *x32;
	uint32_t *u;
	double *rho, *D, *kappa, *kd;  /* n_cls, by class id                */
	double *kappa_p;               /* kappa by single-problem tile
	                                  position (the tile passes)        */
	int kappa_tiled;               /* the last adjoint ran on the tiles:
	                                  kappa_p is current, kappa is not  */
	double *part, *scal;           /* 3 * max(HEMK_NORM_BLOCKS, tiles), 2 */
	double *tot;                   /* 3 per sample                      */
	/* standard errors: tile-order constants and K-wide CG state */
	uint32_t K;                    /* capacity in directions            */
	double *piq, *xq, *aq, *mq, *fdq, *kq;
	uint32_t *jq;
	double *pj, *X, *R, *P, *Q;
	double *prz, *prr, *ppv, *ppfp;
	double *rz0, *rz1, *rr, *bb, *pv, *pfp, *pix, *se, *resid;
} hem_grad_ws;

struct hem_problem {
	pthread_mutex_t mu;
	uint32_t n_tx, n_cls;
	uint64_t nnz;
	hem_tile_config cfg;     /* tile targets and batched chunk */
	uint64_t dev_bytes;
	int group_cls, group_tx;
	void *stream;
	void *ev0, *ev1;

	/* structure (device) */
	uint32_t *d_cls_off, *d_cls_ref, *d_tx_off, *d_tx_ref, *d_len;
	float *d_cnt32;
	double *d_cnt64;
	double *d_nd;      /* n_cls pairs {n_c, D_c} of the FP64 class pass */
	hemk_prob dp;

	/* one solver instance (device) */
	double *d_A, *d_B, *d_C, *d_D;
	float *d_A32, *d_B32, *d_D32;
	void *d_rho;
	uint32_t *d_unique;
	/* plain solve over component tiles: unique counts and class counts
	 * in tile order (state vectors A, B, C are then in tile order too) */
	uint32_t *d_uq;
	double *d_cntq;
	int cntq_ok;               /* d_cntq holds the current class counts */
	double *d_fbpart;  /* fallback: [tile] {ell difference, magnitudes} */
	double *d_tpart;   /* tree norms: [tile][4] partials of r and v      */
	double *d_part1, *d_part2, *d_out;
	/* exact sums by speculation: r/v rows [2][T], scratch per chain
	 * (two chains), and the sums (pass 1 at [0..1], pass 2 at [2..3]) */
	uint32_t spec_L, spec_P;
	double *d_sx, *d_scsum, *d_sdrift, *d_stab, *d_sres;
	double *d_scsq;             /* pass-1 chunk sums of x * x         */
	long long *d_sbase;
	uint8_t *d_snx;
	/* chunk programs of the exact sums (two chains) */
	double *d_swa, *d_swlh;
	uint32_t *d_swrg, *d_swcnt;
	unsigned long long *d_scpub;
	double *d_cls_w, *d_tx_w;  /* member weights (weighted tables)   */
	int weighted;
	hemk_ctl *d_ctl;
	hemk_ctl *h_ctl; /* pinned */
	uint32_t *h_done; /* pinned, set by the device: [0] the loop's done flag,
	                     [1] iterations completed */

	/* component tiles (n_tiles == 0: some component is too large) for
	 * the batched and standard-error kernels */
	hemk_tiles dt;
	uint32_t *h_cls_orig;      /* tile position -> class id (host)      */
	uint32_t *d_cls_orig;      /* same, device                          */
	uint32_t *d_tile_mem[HEM_TILE_ARRAYS]; /* device arrays behind dt   */
	/* the single-problem tile kernels' tiling (larger tiles; one-sample
	 * handles only; n_tiles == 0 otherwise) */
	hemk_tiles dtp;
	uint32_t *d_cls_orig_p;    /* its tile position -> class id         */
	uint32_t *d_ptile_mem[HEM_TILE_ARRAYS];
	uint64_t ptile_rec_bytes;  /* its tile records (algorithmic bytes)  */
	uint32_t *d_ptinfo;        /* per sample {tx0, n_tx, first tile,
	                              tiles} of this tiling                 */
	int counts_integral;       /* class counts are integers < 2^32      */
	double *h_cls_count;       /* host copy of the class counts         */
	hem_batch_ws bw;
	hem_grad_ws gw;
	/* hera-EM's simplified model (hem_solve_hera_em): device state,
	 * allocated on the first call */
	struct {
		int ready;
		uint32_t *occ2mem, *counts;
		uint8_t *lock, *alive, *dead;
		double *P0, *bank, *total, *brow, *zrow, *urow, *lrow, *sums;
		double *p1, *p2, *p3, *acc;
	} he;
	void *ev2, *ev3;

	/* batch graphs superseded by a workspace reallocation or a full cache:
	 * never launched again, destroyed with the handle */
	void **retired;
	int n_retired, cap_retired;

	/* samples: ns independent tables side by side (ns == 1: one table);
	 * h_tx0/h_cls0 [ns+1] first transcript / class of each sample;
	 * device: sinfo [ns][4] {tx0, n_tx, tile0, n_tiles}, tile -> sample,
	 * transcript -> sample */
	uint32_t ns;
	uint32_t *h_tx0, *h_cls0;
	uint32_t *d_sinfo, *d_tsamp, *d_txsamp;

	/* captured iteration graphs and the options each was captured for;
	 * kept until the handle is destroyed */
	void *graph[HEM_GRAPH_CACHE];
	int graph_key[HEM_GRAPH_CACHE];
	int n_graph;
};

/* Tile limits: a component larger than this keeps to the two-pass
 * global kernels (u8 member indices bound the transcripts per tile). */
#define HEM_TILE_TX  64
#define HEM_TILE_CLS 320
/* Components are packed into tiles up to these sizes (a larger component,
 * within the limits above, gets a tile of its own).  The batched and
 * standard-error kernels hold a chunk of replicates per tile in LDS, so
 * their tiles are small and many workgroups share a CU; the single-problem
 * tile kernels (256 threads per tile) run fastest on tiles about four times
 * larger, whose ordered-sum loops are shared by more transcripts and
 * classes. */
#define HEM_TILE_TX_TARGET  24
#define HEM_TILE_CLS_TARGET 120
#define HEM_TILE_NNZ_TARGET 576
#define HEM_PTILE_TX_TARGET  64
#define HEM_PTILE_CLS_TARGET 320
#define HEM_PTILE_NNZ_TARGET 1024
/* a tile holds at most HEM_TILE_CLS transcripts or classes (tiles.c sorts
 * them in buffers of that size) */
_Static_assert(HEM_TILE_TX <= HEM_TILE_CLS &&
               HEM_TILE_TX_TARGET <= HEM_TILE_CLS &&
               HEM_TILE_CLS_TARGET <= HEM_TILE_CLS &&
               HEM_PTILE_TX_TARGET <= HEM_TILE_CLS &&
               HEM_PTILE_CLS_TARGET <= HEM_TILE_CLS,
               "tile sort buffers hold a tile");
#define HEM_TILE_DESC HEMK_TILE_DESC

typedef struct {
	uint32_t n_tiles, n_tx, n_cls, max_tx, max_cls, max_nnz;
	uint64_t nnz;
	uint32_t *desc, *tx, *len, *coff, *toff, *cls_orig, *txq;
	uint32_t *kmem;
	uint16_t *tcls;    /* class-list entry -> local class            */
	double *wcls, *wocc; /* member weights in member order and in
	                        class-list order (weighted tables), or NULL */
	uint32_t *tsamp;   /* tile -> sample (ns > 1 only)               */
	uint32_t *rec;     /* per-tile records, 16-byte units (desc[7])  */
	uint64_t rec_units;
	uint32_t max_rec;  /* bytes of the largest record                */
} hem_tiles_host;

/* tx0[0..ns]: first transcript of each sample (tx0[ns] = T), or NULL with
 * ns = 1.  Components are packed up to the given targets (HEM_TILE_* or
 * HEM_PTILE_*). */
int hem_tiles_build(uint32_t T, uint32_t C, const uint32_t *cls_off,
                    const uint32_t *cls_ref, const uint32_t *tx_len,
                    const uint32_t *tx_off, const uint32_t *tx0,
                    uint32_t ns, const double *member_weight,
                    uint32_t tx_target, uint32_t cls_target,
                    uint32_t nnz_target, hem_tiles_host *h);
void hem_tiles_free_host(hem_tiles_host *h);
/* Bytes of a tile's record (see tiles.c), a multiple of 16. */
/* Either method of the bitwise left-to-right sums. */
static inline int hem_norm_exact(int norm)
{
	return norm == HEM_NORM_EXACT || norm == HEM_NORM_EXACT_WIN;
}

static inline uint64_t hem_tile_rec_bytes(uint32_t nt, uint32_t nc,
                                          uint32_t nz, int weighted)
{
	uint64_t u16s = (uint64_t)(nc + 1) + (nt + 1) + 2 * (uint64_t)nz;
	u16s += u16s & 1;
	uint64_t b = 2 * u16s + 8 * (uint64_t)nt;
	if (weighted)
		b = (b + 7) / 8 * 8 + 16 * (uint64_t)nz;
	return (b + 15) / 16 * 16;
}

/* kernel kinds for per-kernel timing (hem_kernel_profile) */
enum { HEM_PK_DENOM = 0, HEM_PK_GATHER, HEM_PK_NORM1, HEM_PK_NORM2,
       HEM_PK_EXTRAP, HEM_PK_ITER_END, HEM_PK_TMAP12, HEM_PK_TEX3 };
/* kernel kinds of hem_profile_batch */
enum { HEM_BK_MAP12 = 0, HEM_BK_NORM1, HEM_BK_NORM2, HEM_BK_EX3,
       HEM_BK_ITER_END };
/* Largest dynamic LDS of the plain tile kernels; larger tiles keep to the
 * two-pass kernels. */
#define HEM_PTILE_LDS_MAX (64 * 1024)

/* batch.c */
void hem_batch_free(hem_problem *p);
void hem_batch_destroy_retired(hem_problem *p);
/* grad.c */
void hem_grad_free(hem_problem *p);
/* Uploads a tiling into dt (device arrays in mem[], its class map in
 * *cls_orig, the records' bytes in *rec_bytes when given). */
int hem_tiles_upload(hem_problem *p, const hem_tiles_host *h, hemk_tiles *dt,
                     uint32_t **mem, uint32_t **cls_orig,
                     uint64_t *rec_bytes);

/* error.c */
int hem_fail(int code, const char *fmt, ...);
/* Frees the hera-EM workspace of a handle (hera_em.c). */
void hem_hera_em_free(hem_problem *p);
/* The code of the calling thread's last hem_fail. */
int hem_last_code(void);
int hem_dev_fail(int hip_code, const char *what);
int hem_device_ready(void);
/* The current device's arch name (after hem_device_ready). */
const char *hem_device_arch(void);
/* Replicates per batched workgroup for R replicates: the configured chunk,
 * lowered to the largest whose rows fit the LDS budget (0: none fits). */
int hem_batch_chunk(const hem_problem *p, uint32_t R);
/* Enqueues the gather of the class counts into the single-problem tile
 * order (d_cntq) unless it already holds the current counts. */
int hem_cntq_update(hem_problem *p);

#define HEM_TRY_DEV(call, what)                                   \
	do {                                                      \
		int e_ = (call);                                  \
		if (e_)                                           \
			return hem_dev_fail(e_, (what));          \
	} while (0)

#endif
