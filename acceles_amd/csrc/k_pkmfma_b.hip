// Instantiations of pk_scores with B registers for up to 16 column blocks,
// one query tile per MFMA wavefront in half-rounds (two workgroups per CU;
// one translation unit each, for build parallelism).
#include "k_pkmfma.hpp"

dense_fn get_pk_scores_16(int E) { PK_CASES(16, 1) }
