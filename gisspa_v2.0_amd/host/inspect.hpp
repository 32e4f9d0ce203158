// No-GPU inspection modes of the search executable: the parsed config, LST and Euler inputs,
// and dimensions and checksums of images and stacks as the search reads them.  They link only
// the host readers, so the same code also builds as a stand-alone tool (inspect_main.cpp) under
// the address and undefined-behaviour sanitizers.
#pragma once
#include <string>

namespace gsio {
// Runs the inspection mode named by argv[1] (--parse-dump, --mrc-info, --image-info,
// --stack-info) and returns its exit status; sets handled = false (and returns 0) when argv[1]
// names none of them.
int inspect(int argc, char** argv, bool& handled);
// Prints `<prog>: msg` to stderr and returns 1.
int die(const char* prog, const std::string& msg);
}  // namespace gsio
