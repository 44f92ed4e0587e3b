/* ohp.h - the oracle's own truncating high-precision float.
 *
 * TEST INFRASTRUCTURE ONLY: this file belongs to the CPU oracle.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may build or call it.
 *
 * It restates the arithmetic the reference gets from GMP's mpf_class
 * (complex.h:4,14-17; mandelbrot.cpp uses it throughout).  GMP is an unvendored,
 * unpinned dependency (makefile:24 links -lgmp -lgmpxx; no version is named) and
 * its headers are absent from this image, so the published mpf contract is
 * restated instead: results are truncated toward zero, mpf_get_d truncates toward
 * zero.  The precision is P bits with P = 64 * ((bits + 127) / 64 + 1), i.e. at
 * least what mpf holds for a requested precision of `bits` (mpf keeps prec+1
 * limbs).  Every result is the exact result truncated to P bits.
 *
 * Representation: 32-bit limbs, most significant first.
 *   value = sign * 0.d[0] d[1] ... d[n-1] (base 2^32) * 2^exp,  d[0] >= 2^31.
 * This is written independently of the product's 64-bit-limb type so the two can
 * be checked against each other and against mpmath's round-down arithmetic.
 */
#ifndef NEWMAN_ORACLE_OHP_H
#define NEWMAN_ORACLE_OHP_H

#include <stdint.h>

#define OHP_MAXN 84 /* 32-bit limbs: 2688 bits */

typedef struct {
  int n;    /* precision in 32-bit limbs */
  int sign; /* -1, 0, +1 */
  long exp;
  uint32_t d[OHP_MAXN];
} ohp;

int ohp_limbs_for_bits(long bits); /* 2 * ((bits + 127) / 64 + 1) */
void ohp_zero(ohp* r, int n);
void ohp_copy(ohp* r, const ohp* a);          /* truncate a into r->n limbs */
void ohp_set_d(ohp* r, double v);             /* exact */
double ohp_get_d(const ohp* a);               /* truncating, subnormals kept */
double ohp_get_d_2exp(long* e, const ohp* a); /* mpf_get_d_2exp */
void ohp_add(ohp* r, const ohp* a, const ohp* b);
void ohp_sub(ohp* r, const ohp* a, const ohp* b);
void ohp_mul(ohp* r, const ohp* a, const ohp* b);
void ohp_mul_si(ohp* r, const ohp* a, long v);
void ohp_mul_d(ohp* r, const ohp* a, double v);
void ohp_div_ui(ohp* r, const ohp* a, unsigned long v);
int ohp_cmp(const ohp* a, const ohp* b);
int ohp_set_str(ohp* r, const char* s); /* 0 on success */
void ohp_add_d(ohp* r, const ohp* a, double v);

#endif
