// Instantiations of pk_scores with B registers for up to 32 column blocks
// (one translation unit each, for build parallelism).
#include "k_pkmfma.hpp"

dense_fn get_pk_scores_32(int E) { PK_CASES(32, 0) }
