// AXON CPU oracle command line -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Runs the pinned part of the reference protocol for a key=value config and
// prints the report lines plus one machine-readable PINNED line.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

int main(int argc, char** argv) {
    ora_config c;
    ora_config_parse(argc, (const char* const*)argv, &c);
    char msg[256];
    if (ora_config_check(&c, msg, sizeof msg)) { fprintf(stderr, "invalid config: %s\n", msg); return 2; }
    int threads = 1;
    for (int i = 1; i < argc; ++i)
        if (!strncmp(argv[i], "threads=", 8)) threads = atoi(argv[i] + 8);
    std::vector<char> buf(1 << 20);
    ora_pinned p;
    if (ora_run_protocol(&c, threads, &p, buf.data(), buf.size())) return 3;
    fputs(buf.data(), stdout);
    printf("PINNED within=%.4f recall=%.3f other=%.3f margin=%+.3f sustained=%.3f drive6=%.0f asmjacc=%.5f wstat=%u\n",
           p.within_pct, p.recall75, p.other75, p.margin75, p.sustained, p.drive_tt > 6 ? p.drive_t[6] : 0.0,
           p.asm_jacc, p.wstat_final);
    return 0;
}
