// main.cpp -- `kite` command dispatch (src/main.rs:21-41, src/cli.rs:93-127).
// Errors print "Error: <message>" and exit 1; usage errors exit 2.
#include <csignal>
#include <cstdio>
#include <cstring>
#include <string>

#include "cli.hpp"
#include "commands.hpp"
#include "kitefmt.hpp"

namespace {
void on_sigint(int) { _exit(130); }

int usage() {
    fprintf(stderr,
            "kite:  k-mer indexing and throughput engine (MI355X GPU engine)\n\n"
            "Usage:   kite <COMMAND>\n\n"
            "Commands:\n"
            "  count    Count k-mers from FASTA/FASTQ input and write a .kite index.\n"
            "  dump     Dump k-mer counts from a .kite file into a table.\n"
            "  info     Show metadata and summary information for a .kite file.\n"
            "  histo    Generate a k-mer count histogram table.\n"
            "  merge    Merge multiple .kite files into one output file.\n"
            "  profile  Profile k-mer content across samples into a table.\n"
            "  analyze  Analyze multiple .kite files and write TSV tables.\n");
    return 2;
}
}  // namespace

int main(int argc, char **argv) {
    signal(SIGINT, on_sigint);
    if (argc < 2) return usage();
    std::string cmd = argv[1];
    try {
        if (cmd == "count") return kh::run_count(argc, argv);
        if (cmd == "dump") return kh::run_dump(argc, argv);
        if (cmd == "info") return kh::run_info(argc, argv);
        if (cmd == "histo") return kh::run_histo(argc, argv);
        if (cmd == "merge") return kh::run_merge(argc, argv);
        if (cmd == "profile") return kh::run_profile(argc, argv);
        if (cmd == "analyze") return kh::run_analyze(argc, argv);
        if (cmd == "-h" || cmd == "--help" || cmd == "help") {
            usage();
            return 0;
        }
        if (cmd == "-v" || cmd == "--version") {
            printf("kite 0.7.5 (gfx950 engine)\n");
            return 0;
        }
        fprintf(stderr, "error: unrecognized subcommand '%s'\n", cmd.c_str());
        return 2;
    } catch (const kh::UsageError &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 2;
    } catch (const std::exception &e) {
        fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
