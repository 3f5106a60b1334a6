// CPU oracle, part 3: north-star solve methods with no reference counterpart.
// TEST INFRASTRUCTURE ONLY (see oracle.cpp).
#include <cstdint>
extern "C" {
int orc_version() { return 1; }
}
