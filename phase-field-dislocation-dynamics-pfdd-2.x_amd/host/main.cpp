// main.cpp -- pfdd_hip: runs an in.* deck with the GL path on the GPU.
// Command line as the reference's (PFDD.cpp:61-129): -in/-i deck, -log/-l
// file, -screen/-sc file|none, -echo/-e (accepted); plus -check, which parses
// the deck and prints the derived setup without touching the GPU.
#include <cstdio>
#include <cstring>

#include "pfdd_driver.hpp"

int main(int argc, char **argv)
{
    const char *in = nullptr, *log = "log.pfdd";
    bool check = false;
    for (int i = 1; i < argc; ++i) {
        if ((!strcmp(argv[i], "-in") || !strcmp(argv[i], "-i")) && i + 1 < argc) in = argv[++i];
        else if ((!strcmp(argv[i], "-log") || !strcmp(argv[i], "-l")) && i + 1 < argc) log = argv[++i];
        else if ((!strcmp(argv[i], "-screen") || !strcmp(argv[i], "-sc")) && i + 1 < argc) ++i;
        else if ((!strcmp(argv[i], "-echo") || !strcmp(argv[i], "-e")) && i + 1 < argc) ++i;
        else if (!strcmp(argv[i], "-check")) check = true;
        else {
            fprintf(stderr, "ERROR: Invalid command-line argument %s\n", argv[i]);
            return 1;
        }
    }
    FILE *fp = in ? fopen(in, "r") : stdin;
    if (!fp) {
        fprintf(stderr, "ERROR: Cannot open input script %s\n", in);
        return 1;
    }
    pfdd_host::Input input;
    input.check_only = check;
    if (!check && strcmp(log, "none")) input.output.logfile = fopen(log, "w");
    try {
        input.file(fp);
    } catch (const pfdd_host::DeckError &e) {
        fprintf(stderr, "ERROR: %s\n", e.msg.c_str());
        if (input.output.logfile) fprintf(input.output.logfile, "ERROR: %s\n", e.msg.c_str());
        return 1;
    }
    if (fp != stdin) fclose(fp);
    if (input.output.logfile) fclose(input.output.logfile);
    return 0;
}
