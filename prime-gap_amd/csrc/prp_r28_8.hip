// Instantiation unit: radix-2^28 one-lane variants with digits {60, 64}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<60, 64> registered_r28_8;
}
