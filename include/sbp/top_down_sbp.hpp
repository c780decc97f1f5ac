// The reference's header name top_down_sbp.hpp, for callers that keep their includes: everything is declared in sbp.hpp.
#pragma once
#include "sbp.hpp"
