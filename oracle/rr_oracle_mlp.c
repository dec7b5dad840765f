/*
 * rr_oracle_mlp.c -- FP64 restatement of the learned guide (test infrastructure only).
 */
#include "rr_oracle.h"
