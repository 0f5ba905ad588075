// Instantiation unit: radix-2^28 one-lane variants with digits {78}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<78> registered_r28_11;
}
