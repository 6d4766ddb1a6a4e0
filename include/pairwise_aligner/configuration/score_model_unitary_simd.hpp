// Drop-in for rrahn/pairwise_aligner include/pairwise_aligner/configuration/score_model_unitary_simd.hpp:
// the whole configuration API lives in pairwise_aligner/amd/host_api.hpp.
#pragma once
#include <pairwise_aligner/amd/host_api.hpp>
