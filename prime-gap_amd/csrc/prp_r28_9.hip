// Instantiation unit: radix-2^28 one-lane variants with digits {68, 72}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<68, 72> registered_r28_9;
}
