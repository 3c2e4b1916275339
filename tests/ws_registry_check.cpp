// Host unit check of the look-back workspace registry (no HIP): growth
// retires buffers instead of freeing them, buffers of distinct streams never
// alias, and releasing a stream hands back exactly its own buffers.
#include <cstdio>
#include <cstdlib>
#include <set>

#include "../tacuda_amd/csrc/ws_registry.hpp"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                \
    }                                                              \
  } while (0)

using tca::WsRegistry;

static char pool[64][16];
static int next_buf = 0;
static void* fresh() { return pool[next_buf++]; }

int main() {
  WsRegistry r;
  int s1, s2, p1, p2;  // four distinct stream identities
  const WsRegistry::Key a{0, &s1, 0}, a1{0, &s1, 1}, b{0, &s2, 0};

  // first use: nothing live, minimum size
  CHECK(r.find(a, 100) == nullptr);
  CHECK(r.grow_size(a, 100) == WsRegistry::kMinBytes);
  void* small = fresh();
  r.install(a, small, r.grow_size(a, 100));
  CHECK(r.find(a, 100) == small);
  CHECK(r.find(a, WsRegistry::kMinBytes) == small);

  // a graph captured now holds `small`.  A larger eager call grows the
  // stream's workspace: `small` must be retired, not freed.
  CHECK(r.find(a, WsRegistry::kMinBytes + 1) == nullptr);
  const size_t g = r.grow_size(a, WsRegistry::kMinBytes + 1);
  CHECK(g == 2 * WsRegistry::kMinBytes);  // geometric growth
  void* big = fresh();
  r.install(a, big, g);
  CHECK(r.retired() == 1);
  CHECK(r.find(a, 100) == big);

  // other parts and other streams get their own buffers
  CHECK(r.find(a1, 100) == nullptr);
  CHECK(r.find(b, 100) == nullptr);
  void* pa1 = fresh();
  void* pb = fresh();
  r.install(a1, pa1, r.grow_size(a1, 100));
  r.install(b, pb, r.grow_size(b, 100));
  CHECK(r.find(b, 100) != r.find(a, 100));

  // two plans: each has a private stream, so their workspaces never alias
  const WsRegistry::Key k1{0, &p1, 0}, k2{0, &p2, 0};
  void* q1 = fresh();
  void* q2 = fresh();
  r.install(k1, q1, r.grow_size(k1, 5000));
  r.install(k2, q2, r.grow_size(k2, 5000));
  CHECK(r.find(k1, 5000) == q1 && r.find(k2, 5000) == q2 && q1 != q2);

  // destroying plan 1 releases only its stream's buffer
  auto rel = r.release_stream(0, &p1);
  CHECK(rel.size() == 1 && rel[0].base == q1);
  CHECK(r.find(k2, 5000) == q2);

  // releasing stream 1 returns its live buffers and the retired `small`
  rel = r.release_stream(0, &s1);
  std::set<void*> got;
  for (auto& x : rel) got.insert(x.base);
  CHECK(got == (std::set<void*>{small, big, pa1}));
  CHECK(r.retired() == 0);
  CHECK(r.find(b, 100) == pb);

  // cleanup returns everything left, once
  rel = r.release_all();
  got.clear();
  for (auto& x : rel) got.insert(x.base);
  CHECK(got == (std::set<void*>{pb, q2}));
  CHECK(r.release_all().empty());
  CHECK(r.allocations() == 6);
  std::puts("ws_registry ok");
  return 0;
}
