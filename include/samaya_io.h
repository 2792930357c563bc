/* C ABI of the LP path's file formats: MPS in, MPS and the plain-text solution file out.
 * Plain C, never unwinds; returns SMY_OK (0) or an error code of samaya_pdhg.h.
 *
 * Reference interfaces each entry point replaces (paths relative to the samaya tree):
 *   samaya_read_mps ........................ include/samaya_c.h:31-33
 *   read_mps / read_mps_from_string, ParseError include/samaya/io.hpp:11-28
 *                                            (src/io/mps_reader.cpp)
 *   Model::name / row_names / col_names /
 *   col_type ............................... include/samaya/model.hpp (the fields an MPS file sets)
 *   write_solution (the CLI's --solution) .. apps/cli/main.cpp:137-198
 */
#ifndef SAMAYA_IO_H
#define SAMAYA_IO_H

#include <stddef.h>
#include <stdint.h>

#include "samaya_pdlp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reads a free or fixed MPS file (free tried first; quadratic sections are rejected: they belong
 * to the QP path). On failure *out is NULL, the message (with "line N: " when a line is at
 * fault) goes to errbuf (may be NULL) and SMY_INVALID is returned for malformed input, SMY_ERROR
 * when the file cannot be opened. */
int smy_read_mps(const char* path, smy_model** out, char* errbuf, size_t errlen);
/* The same from text in memory (len bytes, need not be NUL-terminated). threads <= 0: the
 * machine's hardware threads, at most 16; the answer does not depend on it. error_line (may be
 * NULL) receives the failing line number, 0 when the whole model is at fault. */
int smy_read_mps_text(const char* text, size_t len, int threads, smy_model** out, char* errbuf,
                      size_t errlen, long long* error_line);

/* The model's NAME (empty if none). */
int smy_model_name(const smy_model* model, char* buf, size_t len);
/* Row (which = 0) or column (which = 1) names, each followed by '\n', then a NUL; models without
 * names give an empty string. Returns the bytes needed (NUL included); writes only when cap is at
 * least that. */
long long smy_model_names(const smy_model* model, int which, char* buf, size_t cap);
/* Sets the row (which = 0) or column (which = 1) names of a model from `len` bytes of names each
 * followed by '\n' (exactly m or n of them; len 0 clears them), and its integrality marks from n
 * bytes (NULL clears them): what smy_write_mps and smy_write_solution write. */
int smy_model_set_names(smy_model* model, int which, const char* names, size_t len);
int smy_model_set_integrality(smy_model* model, const char* marks);
/* out (n bytes, may be NULL): 1 for a column the file marks integer (INTORG/INTEND markers, BV,
 * LI and UI bounds), else 0. Returns the number of integer columns; the LP path solves the
 * model's relaxation whatever they are. */
long long smy_model_integrality(const smy_model* model, char* out);

/* Free MPS of a model: its names (R0.. / C0.. where it has none or one holds a blank), its
 * integrality as INTORG/INTEND markers, numbers in the shortest decimal form that reads back
 * exactly, infinite values as +-1e30. Reading it back gives the same model, except that a ranged
 * row's far end is recomputed as rhs +- |range| and can differ in its last bit when neither end
 * makes that exact. smy_write_mps_text returns the bytes needed
 * (NUL included) and writes only when cap is at least that. */
int smy_write_mps(const smy_model* model, const char* path, char* errbuf, size_t errlen);
long long smy_write_mps_text(const smy_model* model, char* buf, size_t cap);

/* The --solution file of samaya's CLI for an LP answer: "status <name>", "verified yes|no",
 * "objective <v>" (optimal answers), then "columns n" with one "name value [reduced_cost]" line
 * each and "rows m" with "name activity [dual]", and an "infeasibility_certificate m" or
 * "unbounded_ray n" section; numbers with 17 significant digits. result gives status, verified
 * and objective; column and row lines come only for answered statuses (optimal, infeasible,
 * unbounded) and only with col_value (n); without row_activity
 * (m) it is A col_value; col_dual (n) and row_dual (m) are written for optimal answers when both
 * are given, certificate (m when infeasible, n when unbounded) when given. Names are the
 * model's (c0.. / r0.. where it has none, as the reference writes them). */
int smy_write_solution(const smy_model* model, const char* path, const smy_lp_result* result,
                       const double* col_value, const double* col_dual,
                       const double* row_activity, const double* row_dual,
                       const double* certificate, char* errbuf, size_t errlen);
long long smy_solution_text(const smy_model* model, const smy_lp_result* result,
                            const double* col_value, const double* col_dual,
                            const double* row_activity, const double* row_dual,
                            const double* certificate, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SAMAYA_IO_H */
