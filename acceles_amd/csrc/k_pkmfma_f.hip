// Instantiations of pk_scores with two query tiles per MFMA wavefront and B
// registers for up to 8 column blocks (one translation unit each, for build
// parallelism).
#include "k_pkmfma.hpp"

dense_fn get_pk2_scores_8(int E) { PK_CASES(8, 2) }
