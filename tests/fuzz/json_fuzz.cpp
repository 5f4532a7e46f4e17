// Mutation fuzz of the renderer-config JSON parser (hydrant_amd/csrc/json.hpp):
// seeds are renderer and camera configs; every mutant must parse or raise
// json::Error, never fault.  Built with AddressSanitizer and UBSan by
// tests/test_host_sanitizers.py.
#include <cstdio>
#include <random>
#include <string>
#include "json.hpp"
int main() {
  const std::string seeds[] = {
    R"({"render":{"resolution":[512,512],"renderer":"Volume","params":{"mode":"Volume","density":1e-2,"transfer_fn":{"values":[[0,0,0,0],[1,1,1,1]]}}},"camera":{"orbit":{"arm":[25,65,2.2]}}})",
    R"({"a":"xé\n\"y","b":[true,false,null,-1.5e+300,0.0],"c":{}})",
    R"([1,2,[3,[4,{"k":"v"}]],"\\"])"};
  std::mt19937 rng(1234);
  long ok = 0, err = 0;
  for (int it = 0; it < 60000; ++it) {
    std::string s = seeds[it % 3];
    int n = 1 + rng() % 6;
    for (int k = 0; k < n; ++k) {
      size_t p = s.empty() ? 0 : rng() % s.size();
      switch (rng() % 5) {
        case 0: if (!s.empty()) s[p] = (char)(rng() % 256); break;
        case 1: s.insert(p, 1, "{}[]\",:\\u0e-+.1aZ"[rng() % 18]); break;
        case 2: if (!s.empty()) s.erase(p, 1 + rng() % 4); break;
        case 3: s.resize(p); break;
        case 4: s.insert(p, std::string(rng() % 3 ? 8 : 2000, "[{"[rng() % 2])); break;
      }
    }
    try { hyd::json::parse(s); ++ok; } catch (const std::exception &) { ++err; }
  }
  std::printf("ok %ld err %ld\n", ok, err);
  return 0;
}
