/* CPU restatement of the plane-wave right-hand-side generator -- TEST
 * INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the benchmark's CPU baseline load this
 * (through liborasm.so); the product path never does.
 *
 *   ora_haskgrn_f32  HASKGRN as the reference is compiled: REAL inputs and
 *                    outputs, the single-precision subexpressions of
 *                    haskgrn.f kept in float, everything else in double.
 *   ora_haskgrn_f64  the same under -fdefault-real-8 (every REAL is double).
 *
 * Follows haskgrn.f:31-672 step by step, including two behaviours of the
 * reference that a cleaner formulation would not have:
 *  - a depth at or below the top of the half-space is propagated with the
 *    terms of layer nl-1 (haskgrn.f:431-435 sets i = nl, so ibov = nl-1);
 *  - when the walk enters a new layer, the layer it finishes is restored to
 *    its full-thickness terms with the trig/hyperbolic choice made from the
 *    speeds of layer m, m being the DO variable left over from the previous
 *    depth (haskgrn.f:454,466), and only one layer is restored per entry.
 * Parity: pinned against the compiled reference (oracle/_ref, both
 * precisions) and, for a homogeneous medium, against the closed-form
 * free-surface P and SV responses (tests/test_rhs_oracle.py).
 */
#include <math.h>
#include <stdlib.h>

#define RS float
#define SINRS sinf
#define FN(x) x##_f32
#include "hask_oracle_body.h"
#undef RS
#undef SINRS
#undef FN

#define RS double
#define SINRS sin
#define FN(x) x##_f64
#include "hask_oracle_body.h"
#undef RS
#undef SINRS
#undef FN
