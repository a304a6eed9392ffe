// chooser.hip -- learned parameter chooser (placeholder until the trained
// network is loaded: always defers to the analytic choice).
#include "chooser.h"

namespace dmx {

bool learned_choose(int, const size_t *, size_t, double, double, double, bool,
                    bool, LearnedChoice &) {
  return false;
}

}  // namespace dmx
