// Instantiations of pk_scores with B registers for 24, 30 or 32 column
// blocks and eight MFMA wavefronts of one query tile each per workgroup (one
// translation unit, for build parallelism).
#include "k_pkmfma.hpp"

dense_fn get_pk3_scores_24(int E) { PK_CASES(24, 3) }
dense_fn get_pk3_scores_30(int E) { PK_CASES(30, 3) }
dense_fn get_pk3_scores_32(int E) { PK_CASES(32, 3) }
