// axon -- command-line front end of the MI355X AXON engine.
//
// Drop-in for the reference binary (src/axon.mm main(), mm:669-1911): the
// key=value arguments go to axon_run, which runs the reference's protocol on
// the engine and prints its report to stdout.  Exit status: 0 success, 2 a
// configuration the reference would run into undefined behaviour or a
// snapshot saved with other training keys, 3 any other failure.
#include <cstdio>

#include "axon_api.h"

int main(int argc, char** argv) {
    const axon_status s = axon_run(argc, (const char* const*)argv, nullptr, nullptr);
    if (s == AXON_OK) return 0;
    fprintf(stderr, "axon: %s (%s)\n", axon_last_error(), axon_status_str(s));
    return (s == AXON_ERR_INVALID_CONFIG || s == AXON_ERR_BAD_SNAPSHOT) ? 2 : 3;
}
