/* cdr3_synth.h -- seeded synthetic CDR3 workloads (benchmark inputs only; no
 * counterpart in the reference, whose notebooks read Emerson repertoires). */
#ifndef CDR3_SYNTH_H
#define CDR3_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* n unique sequences; returns total characters written (< 0 on error) */
int64_t xs_generate_nn(uint64_t seed, int32_t n, double mu, double sigma, int32_t lmin, int32_t lmax,
                       double planted_frac, int32_t max_edits, char* chars, int64_t cap_chars,
                       uint32_t* offsets);
/* log-uniform repertoire sizes; returns their sum */
int64_t xs_overlap_sizes(uint64_t seed, int32_t n_reps, int32_t size_min, int32_t size_max,
                         int32_t* sizes);
/* repertoires unique within themselves; shared_frac of each repertoire after
 * the first copies (exactly or with one edit) a sequence of earlier ones;
 * frequencies ~ Zipf(zipf_s) on 1..zipf_max */
int64_t xs_generate_overlap(uint64_t seed, int32_t n_reps, const int32_t* sizes, double shared_frac,
                            double mu, double sigma, int32_t lmin, int32_t lmax, double zipf_s,
                            int32_t zipf_max, char* chars, int64_t cap_chars, uint32_t* offsets,
                            int32_t* freq);
/* a query set: planted_frac of it are one-edit relatives of index rows */
int64_t xs_generate_queries(uint64_t seed, const char* idx_chars, const uint32_t* idx_offsets,
                            int32_t n_index, int32_t n, double mu, double sigma, int32_t lmin,
                            int32_t lmax, double planted_frac, char* chars, int64_t cap_chars,
                            uint32_t* offsets);
/* writes the n sequences as the CLI's input CSV: a header line, then one
   "sequence\n" per row, or "sequence,frequency\n" when freq is not NULL.
   Returns 0, or -1 when the file cannot be written. */
int32_t xs_write_csv(const char* path, const char* chars, const uint32_t* offsets, int32_t n,
                     const int32_t* freq);
#ifdef __cplusplus
}
#endif
#endif
