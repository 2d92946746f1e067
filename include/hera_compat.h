/*
 * hera_compat.h -- hera's own equivalence-class, reference and result layouts.
 *
 * These are the structures the drop-in `estimate_count` reads and returns.
 * They must stay layout-identical to hera's definitions so that an object built
 * against this header links into `hera quant` in place of src/EM.c's object:
 *
 *   Class    src/hash_align.h:99-104   {u32 n; u32 *ref; double count; u32 id}
 *   Gclass   src/hash_align.h:106-110  {u32 n; u32 *map; Class *cls}
 *   Ref_inf  src/hash_align.h:112-121  {n, name_len, ref_name, len, seq, find,
 *                                       eff_len, count}
 *   EM_val   src/EM.h:16-19            {double *grama (TPM); double *overlap
 *                                       (est_counts)}
 *
 * Field names and order follow hera so that the structures are the same type
 * for the C compiler; no hera source is included.
 */
#ifndef HERA_COMPAT_H_
#define HERA_COMPAT_H_

#ifdef __cplusplus
extern "C" {
#endif

/* One equivalence class: the transcripts a multi-mapping fragment hits.
 * Members are u32 transcript ids; they may repeat (paired candidates are
 * concatenated, src/hash_align.c:1604-1618) and need not be sorted.
 * `count` holds an integer number of fragments as a double. */
typedef struct {
	unsigned int n;
	unsigned int *ref;
	double count;
	unsigned int id;
} Class;

/* The class table (global CLASS in hera, src/hash_align.c:35). */
typedef struct {
	unsigned int n;
	unsigned int *map;
	Class *cls;
} Gclass;

/* Per-transcript reference information (global REF_INF, src/hash_align.c:42).
 * The EM reads only n, len (u32 transcript length) and count (unique reads). */
typedef struct {
	unsigned int n;
	unsigned int name_len;
	char *ref_name;
	unsigned int *len;
	char *seq;
	unsigned int *find;
	double *eff_len;
	unsigned int *count;
} Ref_inf;

/* Result of one EM run (src/EM.h:16-19): grama = TPM, overlap = est_counts,
 * both calloc'd so that hera's destroy_emVal / free releases them. */
typedef struct {
	double *grama;
	double *overlap;
} EM_val;

#ifdef __cplusplus
}
#endif
#endif /* HERA_COMPAT_H_ */
