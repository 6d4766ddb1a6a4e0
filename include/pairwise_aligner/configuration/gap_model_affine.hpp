// Drop-in for rrahn/pairwise_aligner include/pairwise_aligner/configuration/gap_model_affine.hpp:
// the whole configuration API lives in pairwise_aligner/amd/host_api.hpp.
#pragma once
#include <pairwise_aligner/amd/host_api.hpp>
