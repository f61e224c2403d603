// FP8 e4m3 (row-scaled) MFMA join instantiations: KC = ceil(d/128) chunks.
#define BFJ_TR TrF8
#define BFJ_SFX f8
#define BFJ_LIST(X) X(1, 16) X(1, 32) X(1, 64) X(1, 96) X(1, 128) X(2, 16) X(2, 32) X(2, 64) X(2, 96) X(2, 128)
#include "join_dispatch.h"
