/* hcx_flags.h — per-path result flag bits of libhcx (included by hcx.h). */
#ifndef HCX_FLAGS_H
#define HCX_FLAGS_H

/* Path flag bits. */
#define HCX_FLAG_TRACKED    0x01u  /* reached t = 1 (reference "converged", kernel ...4pts.cu:258) */
#define HCX_FLAG_INFINITY   0x02u  /* |x|^2 > inf_tol (reference "infinity", :259) */
#define HCX_FLAG_REFINED    0x04u  /* FP64 refinement at t = 1 converged to a finite root */
#define HCX_FLAG_REAL       0x08u  /* refined root has all |Im x_i| <= real_tol */
#define HCX_FLAG_UNIQUE     0x10u  /* first occurrence of this root in its instance (hcx_collect) */
#define HCX_FLAG_BUDGET     0x20u  /* max_steps exhausted before t = 1 */
#define HCX_FLAG_MINSTEP    0x40u  /* dt fell below dt_min */
#define HCX_FLAG_SINGULAR   0x80u  /* zero pivot met in the final refinement */
#define HCX_FLAG_ESCALATED  0x100u /* mixed-precision attempt failed; path re-tracked in FP64 */
#define HCX_FLAG_REAL_REF   0x200u /* reference real test on the tracked root, before refinement:
                                      tracked and all |Im x_i| <= real_ref_tol (Evaluations.cpp:88-110) */
#define HCX_FLAG_UNIQUE_REF 0x400u /* counted unique by the reference's Find_Unique_Sols control flow
                                      (Evaluations.cpp:127-176; set by hcx_collect) */

#endif /* HCX_FLAGS_H */
