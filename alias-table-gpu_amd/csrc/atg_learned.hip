// Learned split index (placeholder).
#include "atg_build.h"
namespace atg {
hipError_t launch_split_learned(const BuildArgs&, hipStream_t) { return hipErrorNotSupported; }
}
