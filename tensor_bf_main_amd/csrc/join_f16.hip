// FP16 MFMA join instantiations: KC = ceil(d/32) chunks, R query rows per wavefront.
#define BFJ_TR TrF16
#define BFJ_SFX f16
#define BFJ_LIST(X) X(1, 16) X(1, 32) X(1, 64) X(1, 128) X(2, 16) X(2, 32) X(2, 64) X(2, 96) X(2, 128) X(3, 32) X(3, 64) X(3, 96) X(3, 128) \
    X(4, 16) X(4, 32) X(4, 64) X(4, 96) X(4, 128) X(5, 32) X(5, 64) X(6, 32) X(6, 64) X(7, 16) X(7, 32) X(7, 64) X(8, 32) X(8, 64)
#include "join_dispatch.h"
