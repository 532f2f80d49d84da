// lt_cli.h — the flag surface of the reference programs (src/Config.h:38-89).
//
// Reference flags, same names and defaults:
//   -f/--file <bcsr>            graph (required)
//   -p/--partition <n>          HBM cache slots            (default: all partitions)
//   -b/--batch <n>              walkers resident at once   (default: all walkers)
//   -w/--walker <n>             walkers                    (default: 2 * num nodes)
//   -s/--partition-size <MB>    partition size with -m     (default: 32)
//   -l/--length <n>             walk length                (default: 10)
//   --runs <n>                  epochs                     (default: 1)
//   --trimmed <n>               parsed and unused, as in the reference
//   --source <v>                ppr source                 (default: 0)
//   --prob <p>                  ppr stop probability       (default: 0.15)
//   -m/--materialized-partition build <file>.part instead of reading it
// Flags added by this engine (they only add behaviour):
//   --scheduler greedy|fifo|learned, --scorer <weights file> (learned),
//   --no-pipeline, --zerocopy-threshold <n>, --span partition|resident,
//   --prefetch-depth <n>, --zc-batch <n>, --device-schedule, --pack,
//   --pack-blocks, --rowptr-resident, --seed <n>, --timers, --json,
//   --queries <file>, --tune (with --queries: autotune the configuration),
//   --resident-batch <n>, --walkers-per-thread <n>, --slice-chunks <n>,
//   --prefetch-plain <0|1|2>, --evict-shortlist <n>, --learned-next <pct>
//   (the lt_config fields of the same names; 0 = the library's default),
//   --no-random-access-reduce, --one-level-queue (append_mode 1 and 2: the
//   reference's figure-12 builds -DNO_RANDOM_ACCESS_REDUCE, -DONE_LEVEL_QUEUE)
//   --split-blocks (with -m: lt_partition_build_blocks, partitions sized for
//   block slots of -s MB), --cache-mb <MB> (without -p: as many cache slots
//   of the table's partition size as fit, the way to size the cache of a
//   --split-blocks table, whose partition size is its largest raw partition)
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sstream>
#include <string>

struct CliConfig {
    std::string file;
    bool has_p = false, has_b = false, has_w = false;
    uint64_t p = 0, b = 0, w = 0;
    uint64_t partition_mb = 32;
    uint32_t length = 10;
    int runs = 1;
    uint32_t trimmed = 0;
    uint32_t source = 0;
    float prob = 0.15f;
    bool materialize = false;
    std::string scheduler = "greedy";
    std::string scorer;      // trained scorer weights (lt_scorer_save) for --scheduler learned
    int prefetch_depth = -1;  // -1: library default
    int zc_batch = 0;         // 0: library default
    int resident_batch = 0;   // 0: library default
    int walkers_per_thread = 0;
    int slice_chunks = 0;
    int prefetch_plain = 0;
    int evict_shortlist = 0;
    int learned_next = 0;
    int append_mode = 0;
    bool device_schedule = false;
    int pack = 0;                // lt_config.pack_partitions: 1 --pack, 2 --pack-blocks
    bool rowptr_resident = false;
    bool split_blocks = false;   // -m: lt_partition_build_blocks instead of lt_partition_build
    uint64_t cache_mb = 0;       // without -p: cache slots = cache_mb MB / partition size
    bool tune = false;           // ppr --queries: pick the configuration with lt_tune_ppr_batch
    bool pipeline = true;
    bool has_zc = false;
    uint64_t zc = 0;
    std::string span = "partition";
    uint64_t seed = 0x4C54524Bull;
    bool timers = false;
    bool json = false;
    std::string queries;
};

template <typename T>
static T cli_num(const char *s) {
    T v{};
    std::istringstream iss(s);
    iss >> v;
    return v;
}

// Accepts "-x v", "--long v" and "--long=v".
static bool cli_parse(int argc, char **argv, CliConfig &c) {
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string val;
        bool has_eq = false;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_eq = true;
        }
        auto next = [&](const char *name) -> const char * {
            if (has_eq) return strdup(val.c_str());
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", name);
                exit(1);
            }
            return argv[++i];
        };
        if (a == "-f" || a == "--file") c.file = next("file");
        else if (a == "-p" || a == "--partition") { c.p = cli_num<uint64_t>(next("partition")); c.has_p = true; }
        else if (a == "-b" || a == "--batch") { c.b = cli_num<uint64_t>(next("batch")); c.has_b = true; }
        else if (a == "-w" || a == "--walker") { c.w = cli_num<uint64_t>(next("walker")); c.has_w = true; }
        else if (a == "-s" || a == "--partition-size") c.partition_mb = cli_num<uint64_t>(next("partition-size"));
        else if (a == "-l" || a == "--length") c.length = cli_num<uint32_t>(next("length"));
        else if (a == "--runs") c.runs = cli_num<int>(next("runs"));
        else if (a == "--trimmed") c.trimmed = cli_num<uint32_t>(next("trimmed"));
        else if (a == "--source") c.source = cli_num<uint32_t>(next("source"));
        else if (a == "--prob") c.prob = cli_num<float>(next("prob"));
        else if (a == "-m" || a == "--materialized-partition") c.materialize = true;
        else if (a == "--scheduler") c.scheduler = next("scheduler");
        else if (a == "--scorer") c.scorer = next("scorer");
        else if (a == "--prefetch-depth") c.prefetch_depth = cli_num<int>(next("prefetch-depth"));
        else if (a == "--zc-batch") c.zc_batch = cli_num<int>(next("zc-batch"));
        else if (a == "--resident-batch") c.resident_batch = cli_num<int>(next("resident-batch"));
        else if (a == "--walkers-per-thread") c.walkers_per_thread = cli_num<int>(next("walkers-per-thread"));
        else if (a == "--slice-chunks") c.slice_chunks = cli_num<int>(next("slice-chunks"));
        else if (a == "--prefetch-plain") c.prefetch_plain = cli_num<int>(next("prefetch-plain"));
        else if (a == "--evict-shortlist") c.evict_shortlist = cli_num<int>(next("evict-shortlist"));
        else if (a == "--learned-next") c.learned_next = cli_num<int>(next("learned-next"));
        else if (a == "--no-random-access-reduce") c.append_mode = 1;
        else if (a == "--one-level-queue") c.append_mode = 2;
        else if (a == "--device-schedule") c.device_schedule = true;
        else if (a == "--pack") c.pack = 1;
        else if (a == "--pack-blocks") c.pack = 2;
        else if (a == "--rowptr-resident") c.rowptr_resident = true;
        else if (a == "--split-blocks") c.split_blocks = true;
        else if (a == "--cache-mb") c.cache_mb = cli_num<uint64_t>(next("cache-mb"));
        else if (a == "--no-pipeline") c.pipeline = false;
        else if (a == "--zerocopy-threshold") { c.zc = cli_num<uint64_t>(next("zerocopy-threshold")); c.has_zc = true; }
        else if (a == "--span") c.span = next("span");
        else if (a == "--seed") c.seed = cli_num<uint64_t>(next("seed"));
        else if (a == "--timers") c.timers = true;
        else if (a == "--json") c.json = true;
        else if (a == "--queries") c.queries = next("queries");
        else if (a == "--tune") c.tune = true;
        else {
            fprintf(stderr, "unknown option %s\n", argv[i]);
            return false;
        }
    }
    return !c.file.empty();
}
