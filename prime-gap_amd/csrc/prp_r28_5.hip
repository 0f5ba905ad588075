// Instantiation unit: radix-2^28 one-lane variants with digits {42, 44}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<42, 44> registered_r28_5;
}
