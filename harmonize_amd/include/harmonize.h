// Entry header of the spec API (reference harmonize/cpp/harmonize.h).
// Specifications and host mains are compiled with `-include harmonize.h`
// (or #include it) and use EventProgram<Spec> / AsyncProgram<Spec>,
// init<P>/exec<P>, util::iter, util::host, util::cli and RemappingBarrier
// exactly as with the reference.
#pragma once

#include "harmonize/preamble.h"
#include "harmonize/wave.h"
#include "harmonize/promise.h"
#include "harmonize/iter.h"
#include "harmonize/host.h"
#include "harmonize/cli.h"
#include "harmonize/program.h"
#include "harmonize/barrier.h"

