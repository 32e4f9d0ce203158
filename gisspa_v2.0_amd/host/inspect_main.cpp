// gisspa_inspect: the no-GPU inspection modes of gisspa_main as a stand-alone program that links
// only the host readers (built with the address and undefined-behaviour sanitizers for the CPU
// tests of the readers).
#include <string>

#include "inspect.hpp"

int main(int argc, char** argv) {
  bool handled = false;
  const int rc = gsio::inspect(argc, argv, handled);
  if (handled) return rc;
  return gsio::die("gisspa_inspect",
                   "usage: gisspa_inspect --parse-dump <config> | --mrc-info <file.mrc> | "
                   "--image-info <file> [index] | --stack-info <file> <count>");
}
