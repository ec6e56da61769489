// Host-only checks of the plugin registry semantics (src/gi_algorithm.h:26-49):
// duplicate names throw std::logic_error, select() of an unknown name throws
// std::runtime_error, select() activates lazily and only once, and
// mi355x_pt validates its knobs.  Makes no device call.
#include <cstdio>
#include <stdexcept>

#include "mi355x_pt.h"

namespace {
int failures = 0;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                             \
    }                                                         \
  } while (0)

struct counting : gi::gi_algorithm {
  int activations = 0, computes = 0;
  explicit counting(const std::string &n) : gi_algorithm(n) {}
  void activate(const gi::scene_desc *s) override {
    ++activations;
    gi_algorithm::activate(s);
  }
  void compute() override { ++computes; }
};
}  // namespace

int main() {
  counting a("alpha"), b("beta");
  bool threw = false;
  try {
    counting dup("alpha");
  } catch (const std::logic_error &) {
    threw = true;
  }
  CHECK(threw);
  threw = false;
  try {
    gi::gi_algorithm::select("gamma");
  } catch (const std::runtime_error &) {
    threw = true;
  }
  CHECK(threw);
  gi::gi_algorithm::select("beta");
  gi::gi_algorithm::select("beta");
  CHECK(gi::gi_algorithm::selected == &b && b.activations == 1 && a.activations == 0);
  gi::gi_algorithm::selected->update();  // non-progressive: update() == compute()
  CHECK(b.computes == 1 && !b.progressive() && !b.in_progress());
  CHECK(gi::gi_algorithm::list().size() == 2);
  {
    gi::mi355x_pt pt(8, 8);
    CHECK(gi::gi_algorithm::get("mi355x_pt") == &pt && pt.progressive() && !pt.in_progress());
    threw = false;
    try {
      pt.path_length(0);
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
      pt.compute();
    } catch (const std::logic_error &) {
      threw = true;
    }
    CHECK(threw);
  }
  CHECK(gi::gi_algorithm::get("mi355x_pt") == nullptr);
  std::printf(failures ? "registry test FAILED\n" : "registry test ok\n");
  return failures ? 1 : 0;
}
