// HPCdaligner: the block-pair plan emitter in front of daligner
// (dalign/HPCdaligner.c).  Writes the daligner jobs that compare every pair
// of blocks of a DB (or of a block range), a few B blocks per call, followed
// by one LAmerge job per block.  Same options and the same text, so existing
// pipelines that run the plan keep working; the daligner lines it emits run
// the GPU daligner (marvel_amd/bin/daligner) unchanged.
//
// Output details kept from the reference:
//   * block names are <dir><root>.<i> when the DB path has a directory part
//     longer than one character (no separator is inserted,
//     dalign/HPCdaligner.c:693-696) and <root>.<i> otherwise;
//   * -M is written as "-s<mem>" (dalign/HPCdaligner.c:683-684, :757-758);
//   * the "# Daligner jobs (n)" count follows the reference's formula, which
//     differs from the number of lines written under -X
//     (dalign/HPCdaligner.c:583-603 against :636-705);
//   * headers are written only when the plan goes to stdout (no -o).
#include <getopt.h>
#include <strings.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Opts {
    int dal_jobs = 4, dal_diag_jobs = 0, merge_jobs = 8;
    int verbose = 0, bias = 0, keep = 0, sort = 0, consecutive = 0, no_trace = 0;
    int asym = 0, identity = 0, ignore = 0;
    int k = 14, w = 6, h = 35, t = 0, hgap = 0, s = 100, l = 1000, run_id = 1, mem = 0, nthreads = 4;
    double erel = 0., mrel = 0.;
    std::string host;
    bool have_host = false;
    int port = -1;
    std::vector<std::string> tracks;
    std::string db, root, dir;
    int fblock = 1, lblock = 1;
    FILE* dal = stdout;
    FILE* mrg = stdout;
};

[[noreturn]] void die(const char* msg)
{
    fprintf(stderr, "%s", msg);
    exit(1);
}

void usage(const char* prog)
{
    fprintf(stderr,
            "\nUsage:\t%s\t [-vbAIKSXT] [-k<int(14)>] [-w<int(6)>] [-h<int(35)>] [-t<int>] [-H<int>]\n"
            " [-M<int>] [-e<double(.70)] [-l<int(1000)>] [-r<int>] [-s<int(100)>]\n"
            " [--dal<int(4)>] [--dalDiag<int(1)>] [--mrg<int(8)>] [-D host[:port]]\n"
            " [-o fileSuffix] [-j<int(4)>] [-mtrack]+  <path:db> [<block:int>[-<range:int>]\n\n",
            prog);
}

// number of blocks from the .db stub: "files = n", n file lines, "blocks = m"
int db_blocks(const std::string& db)
{
    FILE* f = fopen(db.c_str(), "r");
    if (f == nullptr) {
        fprintf(stderr, "failed to open database\n");
        return -1;
    }
    int nfiles = 0, nblocks = 0;
    if (fscanf(f, "files = %d\n", &nfiles) != 1) {
        fclose(f);
        fprintf(stderr, "format error in database file\n");
        return -1;
    }
    std::vector<char> line(30001);
    for (int i = 0; i < nfiles; i++)
        if (fgets(line.data(), 30000, f) == nullptr) {
            fclose(f);
            fprintf(stderr, "format error in database file\n");
            return -1;
        }
    if (fscanf(f, "blocks = %d\n", &nblocks) != 1) {
        fclose(f);
        fprintf(stderr, "could not locate 'blocks' entry\n");
        return -1;
    }
    fclose(f);
    return nblocks;
}

std::string block_name(const Opts& o, int i)
{
    const std::string stem = o.dir.size() > 1 ? o.dir + o.root : o.root;
    return stem + "." + std::to_string(i);
}

// daligner options in the order the reference writes them; -T and -M swap
// places between the two job orders
std::string dal_options(const Opts& o, bool consecutive)
{
    std::string s;
    char b[64];
    if (o.verbose) s += " -v";
    if (o.bias) s += " -b";
    if (o.asym) s += " -A";
    if (o.identity) s += " -I";
    if (o.ignore) s += " -i";
    if (o.k != 14) s += " -k" + std::to_string(o.k);
    if (o.w != 6) s += " -w" + std::to_string(o.w);
    if (o.h != 35) s += " -h" + std::to_string(o.h);
    if (o.t > 0) s += " -t" + std::to_string(o.t);
    if (o.hgap > 0) s += " -H" + std::to_string(o.hgap);
    if (o.erel > .1) {
        snprintf(b, sizeof b, " -e%g", o.erel);
        s += b;
    }
    if (o.mrel > .1) {
        snprintf(b, sizeof b, " -m%g", o.mrel);
        s += b;
    }
    if (o.l != 1000) s += " -l" + std::to_string(o.l);
    if (o.s != 100) s += " -s" + std::to_string(o.s);
    const std::string mem = o.mem > 0 ? " -s" + std::to_string(o.mem) : "";
    const std::string notrace = o.no_trace > 0 ? " -T" : "";
    s += consecutive ? mem + notrace : notrace + mem;
    if (o.have_host) s += " -D" + o.host;
    if (o.port > 0) s += ":" + std::to_string(o.port);
    s += " -r" + std::to_string(o.run_id);
    s += " -j" + std::to_string(o.nthreads);
    for (const std::string& m : o.tracks) s += " -m" + m;
    return s;
}

long parse_int(const char* a) { return strtol(a, nullptr, 10); }

Opts parse(int argc, char** argv)
{
    Opts o;
    static const option longs[] = {
        {"help", no_argument, 0, '?'},        {"verbose", no_argument, 0, 'v'},
        {"bias", no_argument, 0, 'b'},        {"keep", no_argument, 0, 'K'},
        {"check", no_argument, 0, 'C'},       {"asymm", no_argument, 0, 'A'},
        {"identity", no_argument, 0, 'I'},    {"ignore", no_argument, 0, 'i'},
        {"kmer", required_argument, 0, 'k'},  {"bwidth", required_argument, 0, 'w'},
        {"hits", required_argument, 0, 'h'},  {"htimes", required_argument, 0, 't'},
        {"rlen", required_argument, 0, 'H'},  {"mem", required_argument, 0, 'M'},
        {"cor", required_argument, 0, 'e'},   {"alen", required_argument, 0, 'l'},
        {"rid", required_argument, 0, 'r'},   {"trace", required_argument, 0, 's'},
        {"dal", required_argument, 0, 'n'},   {"dalDiag", required_argument, 0, 'N'},
        {"mrg", required_argument, 0, 'c'},   {"dServer", required_argument, 0, 'D'},
        {"out", required_argument, 0, 'o'},   {"nthreads", required_argument, 0, 'j'},
        {"track", required_argument, 0, 'm'}, {"sort", required_argument, 0, 'S'},
        {"jobOrder", no_argument, 0, 'X'},    {"noTrace", no_argument, 0, 'T'},
        {0, 0, 0, 0}};
    for (;;) {
        int idx = 0;
        const int c = getopt_long(argc, argv, "?vbKXTSAIk:w:h:t:H:M:e:l:r:s:n:N:c:D:o:m:j:", longs, &idx);
        if (c == -1) break;
        switch (c) {
        case 'v': o.verbose++; break;
        case 'A': o.asym = 1; break;
        case 'I': o.identity = 1; break;
        case 'b': o.bias = 1; break;
        case 'K': o.keep = 1; break;
        case 'S': o.sort = 1; break;
        case 'X': o.consecutive = 1; break;
        case 'T': o.no_trace = 1; break;
        case 'j': {
            const int n = (int)parse_int(optarg);
            int p = 1;
            while (p < n) p *= 2;
            if (p != n) {
                fprintf(stderr, "invalid number of threads: %d. Must be a power of 2.\n", n);
                exit(1);
            }
            o.nthreads = n;
            break;
        }
        case 'k':
            o.k = (int)parse_int(optarg);
            if (o.k < 4 || o.k > 31) die("Kmer length not accepted! Must be in [4, 31]\n");
            break;
        case 'w': o.w = (int)parse_int(optarg); break;
        case 'h': o.h = (int)parse_int(optarg); break;
        case 't': o.t = (int)parse_int(optarg); break;
        case 'H': o.hgap = (int)parse_int(optarg); break;
        case 'M': o.mem = (int)parse_int(optarg); break;
        case 'e':
            o.erel = strtod(optarg, nullptr);
            if (o.erel < .5 || o.erel >= 1) die("Average correlation not accepted! Must be in [.5, 1)\n");
            break;
        case 'l': o.l = (int)parse_int(optarg); break;
        case 'r':
            o.run_id = (int)parse_int(optarg);
            if (o.run_id < 0 || o.run_id > 999) die("Run id not accepted! Must be in [0, 999]\n");
            break;
        case 's': o.s = (int)parse_int(optarg); break;
        case 'n':
            o.dal_jobs = (int)parse_int(optarg);
            if (o.dal_jobs < 1)
                die("Number of block comparisons per call to daligner not accepted! Must be in > 0! \n");
            break;
        case 'N':
            o.dal_diag_jobs = (int)parse_int(optarg);
            if (o.dal_diag_jobs < 1)
                die("Number of block comparisons per call to daligner not accepted! Must be in > 0! \n");
            break;
        case 'c':
            o.merge_jobs = (int)parse_int(optarg);
            if (o.merge_jobs < 3)
                die("Maximum number of files that will be merged in a single LAmerge command not accepted! "
                    "Must be in > 3! \n");
            break;
        case 'D': {
            const char* colon = strchr(optarg, ':');
            o.have_host = true;
            if (colon) {
                o.port = (int)parse_int(colon + 1);
                o.host.assign(optarg, colon - optarg);
            } else
                o.host = optarg;
            break;
        }
        case 'o': {
            const std::string d = std::string(optarg) + ".dalign.plan", m = std::string(optarg) + ".merge.plan";
            if ((o.dal = fopen(d.c_str(), "w")) == nullptr) {
                fprintf(stderr, "ERROR - Cannot open file %s for writing\n", d.c_str());
                exit(1);
            }
            if ((o.mrg = fopen(m.c_str(), "w")) == nullptr) {
                fprintf(stderr, "ERROR - Cannot open file %s for writing\n", m.c_str());
                exit(1);
            }
            break;
        }
        case 'm': o.tracks.push_back(optarg); break;
        case '?': usage(argv[0]); exit(1);
        default:
            fprintf(stderr, "[ERROR] Unsupported argument: %s\n", argv[optind < argc ? optind : argc - 1]);
            usage(argv[0]);
            exit(1);
        }
    }
    if (optind + 1 > argc) {
        fprintf(stderr, "Required arument: <Database>\n");
        usage(argv[0]);
        exit(1);
    }
    // DB path, with ".db" appended unless it ends in it (any case)
    o.db = argv[optind++];
    if (o.db.size() < 3 || strcasecmp(o.db.c_str() + o.db.size() - 3, ".db") != 0) o.db += ".db";
    const size_t slash = o.db.rfind('/');
    o.dir = slash == std::string::npos ? "." : o.db.substr(0, slash);
    o.root = slash == std::string::npos ? o.db : o.db.substr(slash + 1);
    if (o.root.size() > 3 && strcasecmp(o.root.c_str() + o.root.size() - 3, ".db") == 0)
        o.root.resize(o.root.size() - 3);
    const int nblocks = db_blocks(o.db);
    if (nblocks < 0) exit(1);
    o.fblock = 1;
    o.lblock = nblocks;
    if (optind < argc) {
        const char* arg = argv[optind];
        const char* dash = strchr(arg, '-');
        int f, l;
        if (dash) {
            l = (int)parse_int(dash + 1);
            f = (int)parse_int(std::string(arg, dash - arg).c_str());
        } else
            f = l = (int)parse_int(arg);
        if (f < 1 || f > o.lblock) {
            fprintf(stderr, "Ivalid first block id %d! \n", f);
            exit(1);
        }
        if (l < 1 || l < f || l > o.lblock) {
            fprintf(stderr, "Ivalid last block id %d! \n", l);
            exit(1);
        }
        o.fblock = f;
        o.lblock = l;
    }
    if (o.dal_diag_jobs == 0) o.dal_diag_jobs = o.have_host ? 1 : o.dal_jobs;
    return o;
}

// the reference's job count for the header line
int header_jobs(const Opts& o)
{
    int n = 0;
    if (o.consecutive) {
        for (int i = o.fblock; i <= o.lblock; i++)
            for (int low = o.asym ? o.fblock : i; low < o.lblock;) {
                for (int c = 0; low <= o.lblock && c < o.dal_jobs; c++) low++;
                n++;
            }
    } else {
        for (int i = o.fblock; i <= o.lblock; i++)
            n += 1 + (int)ceil(1. * std::max(i - o.fblock + 1 - o.dal_diag_jobs, 0) / o.dal_jobs);
    }
    return n;
}

void emit_daligner(const Opts& o)
{
    if (o.dal == stdout) fprintf(o.dal, "# Daligner jobs (%d)\n", header_jobs(o));
    if (o.consecutive) {
        // block i against blocks i.. (or all blocks with -A), dal_jobs per call
        const std::string opt = dal_options(o, true);
        for (int i = o.fblock; i <= o.lblock; i++)
            for (int low = o.asym ? o.fblock : i; low <= o.lblock;) {
                std::string line = "daligner" + opt + " " + block_name(o, i);
                for (int c = 0; low <= o.lblock && c < o.dal_jobs; c++, low++) line += " " + block_name(o, low);
                fprintf(o.dal, "%s\n", line.c_str());
            }
        return;
    }
    // rounds over the blocks: the first round pairs each block with itself
    // and the dal_diag_jobs - 1 blocks below it, later rounds take the next
    // dal_jobs blocks further down, until every pair (i >= j) is covered
    const std::string opt = dal_options(o, false);
    const int n = o.lblock - o.fblock + 1;
    const long pairs = (long)n * (n + 1) / 2;
    long done = 0;
    for (int round = 0; done < pairs; round++) {
        for (int i = o.fblock; i <= o.lblock; i++) {
            int hgh, low;
            if (round == 0) {
                hgh = i;
                low = std::max(o.fblock, hgh - o.dal_diag_jobs + 1);
            } else {
                hgh = i - (o.dal_diag_jobs + (round - 1) * o.dal_jobs);
                low = std::max(o.fblock, hgh - o.dal_jobs + 1);
            }
            if (hgh < low) continue;
            std::string line = "daligner" + opt + " " + block_name(o, i);
            for (int j = hgh; j >= low; j--, done++) line += " " + block_name(o, j);
            fprintf(o.dal, "%s\n", line.c_str());
        }
        if (round == 0 && o.dal == stdout && o.have_host) fprintf(o.dal, "# end of diagonal\n");
    }
}

void emit_merge(const Opts& o)
{
    if (o.dal == stdout) fprintf(o.mrg, "# merge jobs(%d)\n", o.lblock - o.fblock + 1);
    for (int i = o.fblock; i <= o.lblock; i++) {
        std::string line = "LAmerge";
        if (o.verbose) line += " -" + std::string(o.verbose, 'v');
        if (o.keep) line += " -k";
        if (o.sort) line += " -s";
        line += " -n " + std::to_string(o.merge_jobs);
        char dir[32];
        snprintf(dir, sizeof dir, "%03d_%05d", o.run_id, i);
        fprintf(o.mrg, "%s %s %s.%d.las %s \n", line.c_str(), o.db.c_str(), o.root.c_str(), i, dir);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const Opts o = parse(argc, argv);
    emit_daligner(o);
    emit_merge(o);
    if (o.dal != stdout) fclose(o.dal);
    if (o.mrg != stdout) fclose(o.mrg);
    return 0;
}
