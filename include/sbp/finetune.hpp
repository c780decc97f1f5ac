// The reference's header name finetune.hpp, for callers that keep their includes: everything is declared in sbp.hpp.
#pragma once
#include "sbp.hpp"
