// hj_strat.hpp -- parallel Jacobi strategies: generators and the strat.so loader.
#ifndef HJ_STRAT_HPP
#define HJ_STRAT_HPP

#include <cstdint>
#include <vector>

namespace hj {

// One strategy: steps x pairs x 2 indices (p < q), flattened row-major.
struct Strategy {
  unsigned n = 0, steps = 0, pairs = 0;
  std::vector<uint16_t> tab;  // [steps][pairs][2]
};

// Family names in the order of their IDs 1..6 (share/HypJacL.hpp:118-152).
extern const char *const kFamilies[6];

// Step count of a family of order n (share/HypJacL.cu:24-81): mmstep has n
// steps, the cyclic families n-1.  Returns 0 for an unknown family.
unsigned family_steps(const char *fam, unsigned n);
unsigned family_id(const char *fam);

// True if this library can generate the family at order n (the sizes whose
// construction is a closed form; see the generators).
bool can_generate(const char *fam, unsigned n);

// Generate the table; returns false if !can_generate.
bool generate(const char *fam, unsigned n, Strategy &out);

}  // namespace hj

#endif
