// initial.hpp -- cfg::initial (configuration/initial.hpp:74 in the reference): the rule a
// configuration chain starts from; defined in the MI355X host layer.
#pragma once
#include <pairwise_aligner/amd/host_api.hpp>
