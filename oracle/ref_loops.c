/* Array loops over the reference's scalar C ABI, for timing the reference's
 * own CPU path (bench.py's cpu_baseline; test infrastructure, never part of
 * the product). The reference has no array-level entry points for these
 * functions: its Fortran generics are elemental loops over the by-value
 * calls expdd, logdd, sindd, erfdd, sqrtdd (include/multifloats/float64x2.h:
 * 453-529, fsrc/multifloats.fypp:709-736), and its Fortran + - * / are
 * native; the loops below make the same calls through the C ABI
 * (adddd ... divdd, float64x2.h:396-410). Linked against oracle/_ref/
 * libmfref.so, the reference's own math translation unit. */
#include <stdint.h>

typedef struct {
  double hi, lo;
} f2;

f2 adddd(f2, f2);
f2 subdd(f2, f2);
f2 muldd(f2, f2);
f2 divdd(f2, f2);
f2 sqrtdd(f2);
f2 expdd(f2);
f2 logdd(f2);
f2 sindd(f2);
f2 erfdd(f2);
f2 erfcdd(f2);
f2 erfcxdd(f2);
f2 atandd(f2);
f2 asindd(f2);
f2 acosdd(f2);

/* op: 0 sqrt, 1 exp, 2 log, 3 sin, 4 erf (the C ABI's MFHIP_* unary codes),
 * and 16 erfc, 17 erfcx, 29 atan, 31 asin, 32 acos (likewise) */
int refloop_unary(int op, const f2 *a, f2 *c, int64_t n) {
  f2 (*fn)(f2) = 0;
  switch (op) {
    case 0: fn = sqrtdd; break;
    case 1: fn = expdd; break;
    case 2: fn = logdd; break;
    case 3: fn = sindd; break;
    case 4: fn = erfdd; break;
    case 16: fn = erfcdd; break;
    case 17: fn = erfcxdd; break;
    case 29: fn = atandd; break;
    case 31: fn = asindd; break;
    case 32: fn = acosdd; break;
    default: return -1;
  }
  for (int64_t i = 0; i < n; ++i) c[i] = fn(a[i]);
  return 0;
}

/* op: 0 add, 1 sub, 2 mul, 3 div (the C ABI's MFHIP_* binary codes) */
int refloop_binary(int op, const f2 *a, const f2 *b, f2 *c, int64_t n) {
  f2 (*fn)(f2, f2) = 0;
  switch (op) {
    case 0: fn = adddd; break;
    case 1: fn = subdd; break;
    case 2: fn = muldd; break;
    case 3: fn = divdd; break;
    default: return -1;
  }
  for (int64_t i = 0; i < n; ++i) c[i] = fn(a[i], b[i]);
  return 0;
}
