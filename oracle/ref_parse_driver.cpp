// Driver around the reference's own input parsers (src/EMReader/DataReader2.cpp, compiled from
// /root/reference by oracle/ref.mk into oracle/_ref/).  Prints the parsed config, LST and Euler
// data in the canonical format of `gisspa_main --parse-dump`, so tests can compare the two.
// TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <map>
#include <string>
#include <variant>

#include "DataReader2.h"

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  Config conf(argv[1]);
  std::map<std::string, Config::Node> sorted(conf.value.begin(), conf.value.end());
  for (const auto& kv : sorted) {
    if (std::holds_alternative<int>(kv.second))
      std::printf("config %s %d\n", kv.first.c_str(), std::get<int>(kv.second));
    else if (std::holds_alternative<float>(kv.second))
      std::printf("config %s %.6g\n", kv.first.c_str(), std::get<float>(kv.second));
    else
      std::printf("config %s %s\n", kv.first.c_str(), std::get<std::string>(kv.second).c_str());
  }
  for (const auto& e : LST::load(conf.gets("input")))
    std::printf("lst %d %s %.6f %.6f %.6f\n", e.unused, e.rpath.c_str(), e.defocus, e.dfdiff, e.dfang);
  EulerData eu(conf.gets("eulerfile"));
  for (size_t i = 0; i < eu.size(); ++i)
    std::printf("euler %.6g %.6g %.6g\n", eu.euler1[i], eu.euler2[i], eu.euler3[i]);
  return 0;
}
