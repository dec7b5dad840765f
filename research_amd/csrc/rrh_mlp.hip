// rrh_mlp.hip -- learned guide (MLP -> 8-lobe vMF mixture).
#include "rrh_internal.h"

struct rrh_mlp {
    rrh_ctx* ctx = nullptr;
};
