// FP64 MFMA join instantiations: KC = ceil(d/8) chunks (two K=4 MFMA steps each),
// padded up to the next built value.
#define BFJ_TR TrF64
#define BFJ_SFX f64
#define BFJ_LIST(X) X(1, 16) X(1, 32) X(1, 64) X(2, 16) X(2, 32) X(2, 64) X(3, 16) X(3, 32) X(3, 64) \
    X(4, 16) X(4, 32) X(4, 64) X(5, 16) X(5, 32) X(5, 64) X(6, 16) X(6, 32) X(6, 64) X(8, 16) X(8, 32) X(8, 64) \
    X(12, 16) X(12, 32) X(16, 16) X(16, 32) X(20, 16) X(24, 16) X(28, 16) X(32, 16)
#include "join_dispatch.h"
