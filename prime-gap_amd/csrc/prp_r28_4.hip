// Instantiation unit: radix-2^28 one-lane variants with digits {38, 40}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<38, 40> registered_r28_4;
}
