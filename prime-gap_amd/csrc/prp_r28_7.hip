// Instantiation unit: radix-2^28 one-lane variants with digits {52, 56}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<52, 56> registered_r28_7;
}
