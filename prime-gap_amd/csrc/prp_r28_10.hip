// Instantiation unit: radix-2^28 one-lane variants with digits {76}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<76> registered_r28_10;
}
