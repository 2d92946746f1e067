/*
 * writer.c -- hera's abundance tables from an EM result.
 *
 * write_result (src/EM.c:152-205) writes, next to its HDF5 output:
 *   <dir>/<prefix>_abundance.tsv       one row per transcript
 *   <dir>/<prefix>_abundance.gene.tsv  one row per run of consecutive
 *                                      transcripts of the same gene
 * with the row formats of src/EM.c:178-184 and :188-205.  These functions
 * take the arrays the reference reads from its globals (REF_INF names,
 * unique counts, lengths, effective lengths; GENE_MAP gene ids and names)
 * as arguments, so a caller holding an EM result on the host can write
 * the same files.  The HDF5 file (src/EM.c:208-231) is not written.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hem_internal.h"

/* A name block of len bytes printed as %s: up to its first NUL (the
 * reference's blocks are NUL-terminated; a full block is not read past). */
static void put_name(FILE *f, const char *blk, size_t len)
{
	fwrite(blk, 1, strnlen(blk, len), f);
}

static FILE *open_out(const char *path, char **buf)
{
	FILE *f = fopen(path, "w");
	*buf = NULL;
	if (!f)
		return NULL;
	*buf = malloc(1 << 20);
	if (*buf)
		setvbuf(f, *buf, _IOFBF, 1 << 20);
	return f;
}

static int close_out(FILE *f, char *buf, const char *path)
{
	const int bad = ferror(f);
	const int cl = fclose(f);
	free(buf);
	if (bad || cl)
		return hem_fail(HEM_ERR_IO, "writing %s failed", path);
	return HEM_OK;
}

int hem_write_abundance(const char *path, uint32_t n, const char *names,
                        uint32_t name_len, const uint32_t *unique,
                        const uint32_t *len, const double *eff_len,
                        const double *est, const double *tpm)
{
	char *buf;
	FILE *f;
	if (!path || (n && (!names || !name_len || !unique || !len ||
	                    !eff_len || !est || !tpm)))
		return hem_fail(HEM_ERR_INVALID, "null argument");
	if (!(f = open_out(path, &buf)))
		return hem_fail(HEM_ERR_IO, "cannot open %s", path);
	/* src/EM.c:178-184 */
	fputs("#target_id\tunique_map\tlength\teff_length\test_counts\ttpm\n",
	      f);
	for (uint32_t i = 0; i < n; ++i) {
		put_name(f, names + (size_t)i * name_len, name_len);
		fprintf(f, "\t%u\t%u\t%f\t%f\t%f\n", unique[i], len[i],
		        eff_len[i], est[i], tpm[i]);
	}
	return close_out(f, buf, path);
}

int hem_write_gene_abundance(const char *path, uint32_t n, const char *names,
                             uint32_t name_len, const uint32_t *gene,
                             const char *gene_names, uint32_t l_gene,
                             uint32_t n_gene, const double *est,
                             const double *tpm)
{
	char *buf;
	FILE *f;
	if (!path || (n && (!names || !gene || !gene_names || !est || !tpm)))
		return hem_fail(HEM_ERR_INVALID, "null argument");
	/* both columns start 16 bytes into their name blocks */
	if (n && (name_len <= 16 || l_gene <= 16))
		return hem_fail(HEM_ERR_INVALID, "name blocks must be longer "
		                                 "than 16 bytes");
	for (uint32_t i = 0; i < n; ++i)
		if (gene[i] >= n_gene)
			return hem_fail(HEM_ERR_INVALID,
			                "gene id %u of transcript %u >= %u",
			                gene[i], i, n_gene);
	if (!(f = open_out(path, &buf)))
		return hem_fail(HEM_ERR_IO, "cannot open %s", path);
	/* src/EM.c:187-205: the run's sums accumulate in float; the row is
	 * named by the run's last transcript.  (The reference compares
	 * gene[i + 1] also at the last transcript, one past its array; here
	 * the end of the array ends the run.) */
	fputs("#gene_id\tgene_name\test_counts\ttpm\n", f);
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t g = gene[i];
		float s0 = (float)est[i], s1 = (float)tpm[i];
		while (i + 1 < n && gene[i + 1] == g) {
			++i;
			s0 = (float)((double)s0 + est[i]);
			s1 = (float)((double)s1 + tpm[i]);
		}
		put_name(f, names + (size_t)i * name_len + 16, name_len - 16);
		fputc('\t', f);
		put_name(f, gene_names + (size_t)g * l_gene + 16, l_gene - 16);
		fprintf(f, "\t%f\t%f\n", (double)s0, (double)s1);
	}
	return close_out(f, buf, path);
}

/* The reference's file names: dir, a '/' if it has none at its end, then
 * prefix and '_' when prefix is not empty (src/EM.c:161-175). */
static char *out_path(const char *dir, const char *prefix, const char *name)
{
	const size_t l = strlen(dir), p = strlen(prefix);
	char *s = malloc(l + p + strlen(name) + 3);
	if (!s)
		return NULL;
	size_t k = 0;
	memcpy(s, dir, l);
	k = l;
	if (l == 0 || dir[l - 1] != '/')
		s[k++] = '/';
	memcpy(s + k, prefix, p);
	k += p;
	if (p > 0)
		s[k++] = '_';
	strcpy(s + k, name);
	return s;
}

int hem_write_result(const char *out_dir, const char *prefix, uint32_t n,
                     const char *names, uint32_t name_len,
                     const uint32_t *unique, const uint32_t *len,
                     const double *eff_len, const uint32_t *gene,
                     const char *gene_names, uint32_t l_gene,
                     uint32_t n_gene, const double *est, const double *tpm)
{
	char *a, *b;
	int rc;
	if (!out_dir || !prefix)
		return hem_fail(HEM_ERR_INVALID, "null argument");
	a = out_path(out_dir, prefix, "abundance.tsv");
	b = out_path(out_dir, prefix, "abundance.gene.tsv");
	if (!a || !b) {
		free(a);
		free(b);
		return hem_fail(HEM_ERR_NOMEM, "path allocation");
	}
	rc = hem_write_abundance(a, n, names, name_len, unique, len, eff_len,
	                         est, tpm);
	if (!rc && gene)
		rc = hem_write_gene_abundance(b, n, names, name_len, gene,
		                              gene_names, l_gene, n_gene, est,
		                              tpm);
	free(a);
	free(b);
	return rc;
}
