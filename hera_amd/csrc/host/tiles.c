/*
 * tiles.c -- connected components of the class graph, packed into LDS tiles.
 *
 * hera's classes never span unrelated transcripts (a class is the candidate
 * set of one fragment, src/hash_align.c:1564-1602), so the class-transcript
 * graph falls apart into small connected components.  One EM map touches a
 * component's transcripts and classes only, so a workgroup that holds a whole
 * component in LDS can run the class pass and the transcript pass without
 * leaving the CU.  Components are packed, in order of their smallest
 * transcript id, into tiles of about HEM_TILE_TX_TARGET transcripts,
 * HEM_TILE_CLS_TARGET classes and HEM_TILE_NNZ_TARGET members; a component
 * above HEM_TILE_TX transcripts or HEM_TILE_CLS classes is not tiled.
 *
 * Per tile the device holds:
 *   desc     {tx_base, n_tx, cls_base, n_cls, member base, members,
 *            class-list base, 0}
 *   txq      transcript id -> tile position             (all transcripts)
 *   tx       global transcript ids, within a tile by descending
 *            occurrence count (then ascending id)     (n_tx)
 *   len      their lengths                             (n_tx)
 *   coff     member offsets of the tile's classes      (n_cls + 1 shared)
 *   kmem     members in hera's order, each (local class << 8) | local
 *            transcript                                 (members)
 *   toff     class-list offsets of the tile's tx        (n_tx + 1 shared)
 *   tcls     per transcript, the local class (u16) of each of its
 *            occurrences, in ascending original class id and, within a
 *            class, member order -- the order hera's class loop adds into
 *            overlap[t] (src/EM.c:276-285)
 * Classes are renumbered in tile order ("tile position"), within a tile by
 * descending member count (then ascending id); per-replicate class counts
 * are stored in that order so a tile reads one contiguous block.  cls_orig
 * maps a tile position back to the caller's class id.  Local indices are
 * labels only: every sum runs in the order its list gives (members in
 * hera's order, each transcript's class list in ascending original class
 * id).  Ordering by length puts entities of similar length in the same
 * wave of the batched kernels, whose waves loop to their longest entity.
 * Classes without members never contribute (D_c = 0) and are dropped.
 */
#include <stdlib.h>
#include <string.h>

#include "hem_internal.h"

static int cmp_u64(const void *a, const void *b)
{
	const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return x < y ? -1 : x > y;
}

/* Orders ids[0..n) by descending len(id), then ascending id. */
static void order_by_length(uint32_t *ids, uint32_t n, const uint32_t *off,
                            uint64_t *key)
{
	for (uint32_t i = 0; i < n; ++i)
		key[i] = ((uint64_t)(0xFFFFFFFFu - (off[ids[i] + 1] -
		                                    off[ids[i]])) << 32) | ids[i];
	qsort(key, n, sizeof(uint64_t), cmp_u64);
	for (uint32_t i = 0; i < n; ++i)
		ids[i] = (uint32_t)key[i];
}

static uint32_t uf_find(uint32_t *par, uint32_t x)
{
	while (par[x] != x) {
		par[x] = par[par[x]];
		x = par[x];
	}
	return x;
}

/* One contiguous record per tile for the single-problem tile kernels
 * (ptile_kernels.hip), 16-byte aligned, its offset in 16-byte units in
 * desc[7]: u16 class offsets (n_cls + 1) and class-list offsets
 * (n_tx + 1), both relative to the tile; u16 members
 * (local class << 6) | local transcript in class order, then the same u16
 * pairs in class-list order (each transcript's occurrences in the order its
 * sum takes them); then, 4-byte aligned, u32 lengths and u32 transcript
 * ids (n_tx each); for a weighted table then, 8-byte aligned, the FP64
 * member weights in class order and in class-list order.  Runs while tcls
 * still holds member positions. */
static int tiles_records(hem_tiles_host *h)
{
	uint64_t units = 0;
	h->max_rec = 0;
	for (uint32_t i = 0; i < h->n_tiles; ++i) {
		const uint32_t *d = h->desc + (size_t)HEM_TILE_DESC * i;
		const uint64_t b = hem_tile_rec_bytes(d[1], d[3], d[5],
		                                      h->wcls != NULL);
		if (b > h->max_rec)
			h->max_rec = (uint32_t)b;
		units += b / 16;
	}
	if (units > 0xFFFFFFFFu)
		return hem_fail(HEM_ERR_RANGE, "tile records exceed 64 GiB");
	h->rec = calloc(units ? units * 4 : 4, 4);
	if (!h->rec)
		return hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
	h->rec_units = units;
	units = 0;
	for (uint32_t i = 0; i < h->n_tiles; ++i) {
		uint32_t *d = h->desc + (size_t)HEM_TILE_DESC * i;
		const uint32_t nt = d[1], nc = d[3], nz = d[5];
		uint16_t *r16 = (uint16_t *)(h->rec + units * 4);
		uint32_t k = 0;
		for (uint32_t c = 0; c <= nc; ++c)
			r16[k++] = (uint16_t)(h->coff[d[2] + c] - d[4]);
		for (uint32_t t = 0; t <= nt; ++t)
			r16[k++] = (uint16_t)(h->toff[d[0] + t] - d[6]);
		for (uint32_t z = 0; z < nz; ++z) {
			const uint32_t m = h->kmem[d[4] + z];
			r16[k++] = (uint16_t)(((m >> 8) << 6) | (m & 0x3Fu));
		}
		for (uint32_t z = 0; z < nz; ++z) {
			const uint32_t m = h->kmem[d[4] + h->tcls[d[6] + z]];
			r16[k++] = (uint16_t)(((m >> 8) << 6) | (m & 0x3Fu));
		}
		k += k & 1;
		uint32_t *r32 = (uint32_t *)r16 + k / 2;
		for (uint32_t t = 0; t < nt; ++t) {
			r32[t] = h->len[d[0] + t];
			r32[nt + t] = h->tx[d[0] + t];
		}
		if (h->wcls) {
			const size_t b = ((size_t)2 * k + 8 * (size_t)nt + 7) / 8 * 8;
			double *w = (double *)((char *)r16 + b);
			memcpy(w, h->wcls + d[4], (size_t)nz * 8);
			memcpy(w + nz, h->wocc + d[6], (size_t)nz * 8);
		}
		d[7] = (uint32_t)units;
		units += hem_tile_rec_bytes(nt, nc, nz, h->wcls != NULL) / 16;
	}
	return HEM_OK;
}

void hem_tiles_free_host(hem_tiles_host *h)
{
	if (!h)
		return;
	free(h->desc);
	free(h->tx);
	free(h->len);
	free(h->coff);
	free(h->kmem);
	free(h->txq);
	free(h->toff);
	free(h->tcls);
	free(h->wcls);
	free(h->wocc);
	free(h->rec);
	free(h->cls_orig);
	free(h->tsamp);
	memset(h, 0, sizeof(*h));
}

/* Builds the tiling of a CSR class table, given the transcripts' occurrence
 * offsets tx_off (the CSC offsets built by hem_problem_create).  Returns
 * HEM_ERR_RANGE when a component exceeds the tile limits. */
int hem_tiles_build(uint32_t T, uint32_t C, const uint32_t *cls_off,
                    const uint32_t *cls_ref, const uint32_t *tx_len,
                    const uint32_t *tx_off, const uint32_t *tx0,
                    uint32_t ns, const double *member_weight,
                    uint32_t tx_target, uint32_t cls_target,
                    uint32_t nnz_target, hem_tiles_host *h)
{
	uint32_t *par = NULL, *comp = NULL, *ctx = NULL, *ccls = NULL;
	uint32_t *ctile = NULL, *lidx = NULL, *pos = NULL;
	uint32_t *cnz = NULL, *txpos = NULL, *mloc = NULL;
	uint32_t *tile_ntx = NULL, *tile_ncls = NULL;
	uint32_t ncomp = 0, ntile = 0, i, k;
	uint64_t zlive = 0;
	int rc = HEM_OK;

	memset(h, 0, sizeof(*h));
	par = malloc((size_t)T * 4);
	comp = malloc((size_t)T * 4);
	lidx = malloc((size_t)T * 4);
	if (!par || !comp || !lidx) {
		rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
		goto out;
	}
	for (i = 0; i < T; ++i)
		par[i] = i;
	for (i = 0; i < C; ++i) {
		if (cls_off[i + 1] == cls_off[i])
			continue;
		uint32_t a = uf_find(par, cls_ref[cls_off[i]]);
		for (k = cls_off[i] + 1; k < cls_off[i + 1]; ++k) {
			uint32_t b = uf_find(par, cls_ref[k]);
			if (a != b) {
				/* keep the smaller id as root */
				if (b < a) {
					par[a] = b;
					a = b;
				} else {
					par[b] = a;
				}
			}
		}
	}
	/* component ids in order of their smallest transcript */
	for (i = 0; i < T; ++i) {
		uint32_t r = uf_find(par, i);
		comp[i] = (r == i) ? ncomp++ : comp[r];
	}
	ctx = calloc(ncomp, 4);
	ccls = calloc(ncomp, 4);
	cnz = calloc(ncomp, 4);
	ctile = malloc((size_t)ncomp * 4);
	if (!ctx || !ccls || !cnz || !ctile) {
		rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
		goto out;
	}
	for (i = 0; i < T; ++i)
		ctx[comp[i]]++;
	for (i = 0; i < C; ++i)
		if (cls_off[i + 1] > cls_off[i]) {
			ccls[comp[cls_ref[cls_off[i]]]]++;
			cnz[comp[cls_ref[cls_off[i]]]] += cls_off[i + 1] - cls_off[i];
			zlive += cls_off[i + 1] - cls_off[i];
		}
	for (i = 0; i < ncomp; ++i)
		if (ctx[i] > HEM_TILE_TX || ccls[i] > HEM_TILE_CLS) {
			rc = hem_fail(HEM_ERR_RANGE,
			              "component with %u transcripts / %u classes "
			              "exceeds the tile limits (%u / %u)", ctx[i],
			              ccls[i], HEM_TILE_TX, HEM_TILE_CLS);
			goto out;
		}
	/* pack components into tiles; with samples, a tile never holds two
	 * (components are in order of their smallest transcript, so a
	 * sample's components are consecutive and its packing is the one
	 * its own table would get) */
	{
		uint32_t nt = 0, nc = 0, nz = 0, smp = 0, cur = 0;
		uint32_t *root = lidx;   /* smallest transcript of each component */
		for (i = T; i-- > 0;)
			root[comp[i]] = i;
		for (i = 0; i < ncomp; ++i) {
			if (ns > 1)
				while (smp + 1 < ns && root[i] >= tx0[smp + 1])
					++smp;
			if (nt && (smp != cur || nt + ctx[i] > tx_target ||
			           nc + ccls[i] > cls_target ||
			           nz + cnz[i] > nnz_target)) {
				ntile++;
				nt = nc = nz = 0;
			}
			cur = smp;
			ctile[i] = ntile;
			nt += ctx[i];
			nc += ccls[i];
			nz += cnz[i];
		}
		ntile++;
	}
	tile_ntx = calloc(ntile + 1, 4);
	tile_ncls = calloc(ntile + 1, 4);
	h->desc = calloc((size_t)ntile * HEM_TILE_DESC, 4);
	h->tx = malloc((size_t)T * 4);
	h->len = malloc((size_t)T * 4);
	h->toff = malloc(((size_t)T + 1) * 4);
	h->tcls = malloc((zlive ? zlive : 1) * 2);
	pos = malloc(((size_t)ntile + 1) * 4);
	if (!tile_ntx || !tile_ncls || !h->desc || !h->tx || !h->len ||
	    !h->toff || !h->tcls || !pos) {
		rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
		goto out;
	}
	for (i = 0; i < T; ++i)
		tile_ntx[ctile[comp[i]]]++;
	for (i = 0; i < C; ++i)
		if (cls_off[i + 1] > cls_off[i])
			tile_ncls[ctile[comp[cls_ref[cls_off[i]]]]]++;
	/* transcripts in tile order, by descending occurrence count within a
	 * tile */
	{
		uint32_t acc = 0;
		uint64_t key[HEM_TILE_CLS];
		for (i = 0; i < ntile; ++i) {
			pos[i] = acc;
			acc += tile_ntx[i];
		}
		for (i = 0; i < T; ++i) {
			uint32_t tl = ctile[comp[i]];
			h->tx[pos[tl]++] = i;
		}
		for (i = 0, acc = 0; i < ntile; acc += tile_ntx[i++])
			order_by_length(h->tx + acc, tile_ntx[i], tx_off, key);
		for (i = 0; i < T; ++i)
			h->len[i] = tx_len[h->tx[i]];
	}
	/* local index = position minus the tile's first position */
	{
		uint32_t acc = 0;
		for (i = 0; i < ntile; ++i) {
			h->desc[HEM_TILE_DESC * i + 0] = acc;
			h->desc[HEM_TILE_DESC * i + 1] = tile_ntx[i];
			acc += tile_ntx[i];
		}
		for (i = 0; i < ntile; ++i)
			for (k = 0; k < tile_ntx[i]; ++k)
				lidx[h->tx[h->desc[HEM_TILE_DESC * i] + k]] = k;
	}
	/* classes in tile order, by descending member count within a tile */
	h->n_cls = 0;
	for (i = 0; i < ntile; ++i)
		h->n_cls += tile_ncls[i];
	h->cls_orig = malloc((h->n_cls ? h->n_cls : 1) * 4);
	h->coff = malloc(((size_t)h->n_cls + 1) * 4);
	h->kmem = malloc((zlive ? zlive : 1) * 4);
	if (member_weight) {
		h->wcls = malloc((zlive ? zlive : 1) * 8);
		h->wocc = malloc((zlive ? zlive : 1) * 8);
	}
	if (!h->cls_orig || !h->coff || !h->kmem ||
	    (member_weight && (!h->wcls || !h->wocc))) {
		rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
		goto out;
	}
	{
		uint32_t acc = 0;
		for (i = 0; i < ntile; ++i) {
			pos[i] = acc;
			h->desc[HEM_TILE_DESC * i + 2] = acc;
			h->desc[HEM_TILE_DESC * i + 3] = tile_ncls[i];
			acc += tile_ncls[i];
		}
		for (i = 0; i < C; ++i) {
			if (cls_off[i + 1] == cls_off[i])
				continue;
			uint32_t tl = ctile[comp[cls_ref[cls_off[i]]]];
			h->cls_orig[pos[tl]++] = i;
		}
		uint64_t key[HEM_TILE_CLS];
		for (i = 0, acc = 0; i < ntile; acc += tile_ncls[i++])
			order_by_length(h->cls_orig + acc, tile_ncls[i], cls_off,
			                key);
	}
	/* member lists in hera's member order, each member as (local class
	 * << 8) | local transcript; mloc
	 * maps a member of the caller's table to its position here */
	{
		const uint64_t nnz = C ? cls_off[C] : 0;
		mloc = malloc((nnz ? nnz : 1) * 4);
		txpos = malloc((nnz ? nnz : 1) * 4);
		if (!mloc || !txpos) {
			rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
			goto out;
		}
		uint64_t z = 0;
		for (uint32_t tl = 0; tl < ntile; ++tl) {
			const uint32_t q0 = h->desc[HEM_TILE_DESC * tl + 2];
			const uint32_t q1 = q0 + h->desc[HEM_TILE_DESC * tl + 3];
			for (uint32_t q = q0; q < q1; ++q) {
				uint32_t c = h->cls_orig[q];
				h->coff[q] = (uint32_t)z;
				for (k = cls_off[c]; k < cls_off[c + 1]; ++k) {
					mloc[k] = (uint32_t)z;
					if (member_weight)
						h->wcls[z] = member_weight[k];
					h->kmem[z++] = ((q - q0) << 8) |
					               lidx[cls_ref[k]];
				}
			}
		}
		h->coff[h->n_cls] = (uint32_t)z;
	}
	/* each transcript's occurrences as members, in ascending class id and
	 * member order (a counting sort of the member positions), then
	 * relative to the first member of the transcript's tile; the records
	 * read these positions, then they become local classes */
	{
		uint32_t *fill = lidx; /* reused: next free slot per transcript */
		for (i = 0; i < T; ++i)
			fill[i] = tx_off[i];
		for (i = 0; i < C; ++i)
			for (k = cls_off[i]; k < cls_off[i + 1]; ++k)
				txpos[fill[cls_ref[k]]++] = k;
		uint64_t z = 0;
		for (uint32_t tl = 0; tl < ntile; ++tl) {
			const uint32_t q0 = h->desc[HEM_TILE_DESC * tl + 0];
			const uint32_t q1 = q0 + h->desc[HEM_TILE_DESC * tl + 1];
			const uint32_t mb = h->coff[h->desc[HEM_TILE_DESC * tl + 2]];
			for (uint32_t q = q0; q < q1; ++q) {
				uint32_t t = h->tx[q];
				h->toff[q] = (uint32_t)z;
				for (k = tx_off[t]; k < tx_off[t + 1]; ++k) {
					if (member_weight)
						h->wocc[z] =
						    member_weight[txpos[k]];
					h->tcls[z++] =
					    (uint16_t)(mloc[txpos[k]] - mb);
				}
			}
		}
		h->toff[T] = (uint32_t)z;
	}
	for (uint32_t tl = 0; tl < ntile; ++tl) {
		uint32_t *d = h->desc + (size_t)HEM_TILE_DESC * tl;
		d[4] = h->coff[d[2]];
		d[5] = h->coff[d[2] + d[3]] - d[4];
		d[6] = h->toff[d[0]];
	}
	h->txq = malloc((size_t)T * 4);
	if (!h->txq) {
		rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
		goto out;
	}
	for (uint32_t q = 0; q < T; ++q)
		h->txq[h->tx[q]] = q;
	h->n_tiles = ntile;
	h->n_tx = T;
	h->nnz = zlive;
	if (ns > 1) {
		/* each tile's sample, from its first transcript */
		uint32_t smp = 0;
		h->tsamp = malloc((size_t)ntile * 4);
		if (!h->tsamp) {
			rc = hem_fail(HEM_ERR_NOMEM, "tiling: host allocation");
			goto out;
		}
		for (i = 0; i < ntile; ++i) {
			const uint32_t t = h->tx[h->desc[HEM_TILE_DESC * i]];
			while (smp + 1 < ns && t >= tx0[smp + 1])
				++smp;
			h->tsamp[i] = smp;
		}
	}
	h->max_tx = h->max_cls = h->max_nnz = 0;
	for (i = 0; i < ntile; ++i) {
		uint32_t q0 = h->desc[HEM_TILE_DESC * i + 2], q1 = q0 + tile_ncls[i];
		uint32_t z = h->coff[q1] - h->coff[q0];
		if (tile_ntx[i] > h->max_tx)
			h->max_tx = tile_ntx[i];
		if (tile_ncls[i] > h->max_cls)
			h->max_cls = tile_ncls[i];
		if (z > h->max_nnz)
			h->max_nnz = z;
	}
	if ((rc = tiles_records(h)))
		goto out;
	for (i = 0; i < ntile; ++i) {
		const uint32_t *d = h->desc + (size_t)HEM_TILE_DESC * i;
		const uint32_t nz = h->toff[d[0] + d[1]] - d[6];
		for (uint32_t z = d[6]; z < d[6] + nz; ++z)
			h->tcls[z] = (uint16_t)(h->kmem[d[4] + h->tcls[z]] >> 8);
	}
out:
	if (rc)
		hem_tiles_free_host(h);
	free(par);
	free(comp);
	free(ctx);
	free(ccls);
	free(cnz);
	free(ctile);
	free(txpos);
	free(mloc);
	free(lidx);
	free(pos);
	free(tile_ntx);
	free(tile_ncls);
	return rc;
}

int hem_classes_tiling(const hem_classes *c, uint32_t *n_tiles,
                       uint32_t *max_tx, uint32_t *max_cls)
{
	uint32_t *tx_off, i;
	uint64_t nnz;
	hem_tiles_host h;
	int rc;
	if (!c || !c->tx_len || !c->n_tx || (c->n_cls && !c->cls_off))
		return hem_fail(HEM_ERR_INVALID, "null class table");
	if (!n_tiles || !max_tx || !max_cls)
		return hem_fail(HEM_ERR_INVALID, "null output");
	nnz = c->n_cls ? c->cls_off[c->n_cls] : 0;
	for (uint64_t k = 0; k < nnz; ++k)
		if (c->cls_ref[k] >= c->n_tx)
			return hem_fail(HEM_ERR_INVALID, "transcript id out of range");
	tx_off = calloc((size_t)c->n_tx + 1, 4);
	if (!tx_off)
		return hem_fail(HEM_ERR_NOMEM, "host allocation");
	for (uint64_t k = 0; k < nnz; ++k)
		tx_off[c->cls_ref[k] + 1]++;
	for (i = 0; i < c->n_tx; ++i)
		tx_off[i + 1] += tx_off[i];
	rc = hem_tiles_build(c->n_tx, c->n_cls, c->cls_off, c->cls_ref,
	                     c->tx_len, tx_off, NULL, 1, NULL,
	                     HEM_TILE_TX_TARGET, HEM_TILE_CLS_TARGET,
	                     HEM_TILE_NNZ_TARGET, &h);
	free(tx_off);
	if (rc == HEM_ERR_RANGE) {
		*n_tiles = 0;
		*max_tx = *max_cls = 0;
		return HEM_OK;
	}
	if (rc)
		return rc;
	*n_tiles = h.n_tiles;
	*max_tx = h.max_tx;
	*max_cls = h.max_cls;
	hem_tiles_free_host(&h);
	return HEM_OK;
}
