// host/tools_main.cpp -- the reference's offline input tools as command lines
// over libg2host.so; the program's name picks the tool:
//   preprocess -f G.txt|G.bin          -> <G>.bin  (kmeans/preprocess.cpp)
//   kcore -f G -nb K                   -> <G>.bin recoded, <G>.bin.2.hop.vbmap
//                                         (kmeans/kcore.cpp:232-253)
//   kmeans -f G.bin -k 4|4x4 -i 20 [-seed S] [-o prefix]
//                                      -> <prefix><k>, one cluster id per line
//                                         (kmeans/kmeans.cpp:736-794)
//   view_pack -gf G.bin -pf P -h 2 -m GB -s pool -o 0|1 -d 1
//                                      -> <G>.bin.<h>.hop.vbmap and .vbstat
//                                         (kmeans/view_pack.cpp:330-365)
//   toGAMMA G | toGSI G | toPangolin G -> the graph in those systems' input
//                                         formats (converter/*.cpp)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "g2host.h"

namespace {

std::string opt(int argc, char** argv, const char* name, const std::string& def) {
    for (int i = 1; i + 1 < argc; i++)
        if (!std::strcmp(argv[i], name)) return argv[i + 1];
    return def;
}

[[noreturn]] void fail(const std::string& what) {
    std::printf("%s: %s\n", what.c_str(), g2h_last_error());
    std::exit(-1);
}

std::string prefix_of(const std::string& f) { return f.substr(0, f.rfind('.')); }

g2h_graph* load(const std::string& f) {
    g2h_graph* g = nullptr;
    const std::string suffix = f.substr(f.rfind('.') + 1);
    const int rc = suffix == "bin" ? g2h_graph_read_bin(f.c_str(), &g) : g2h_graph_read_snap(f.c_str(), &g);
    if (rc) fail("cannot read " + f);
    std::printf("vertex_count=%llu, edge_count=%llu\n", (unsigned long long)g2h_graph_nv(g),
                (unsigned long long)g2h_graph_ne(g));
    return g;
}

int preprocess(int argc, char** argv) {
    const std::string f = opt(argc, argv, "-f", "./data/com-dblp.ungraph.txt");
    g2h_graph* g = load(f);
    if (g2h_graph_preprocess(g)) fail("preprocess");
    const std::string out = prefix_of(f) + ".bin";
    if (g2h_graph_write_bin(g, out.c_str())) fail("cannot write " + out);
    std::printf("wrote %s\n", out.c_str());
    g2h_graph_free(g);
    return 0;
}

int kcore(int argc, char** argv) {
    const std::string f = opt(argc, argv, "-f", "../data/com-dblp.ungraph.txt");
    const uint32_t nb = (uint32_t)std::atoi(opt(argc, argv, "-nb", "1").c_str());
    if (nb < 1) fail("-nb must be >= 1");
    g2h_graph* g = load(f);
    const uint64_t nv = g2h_graph_nv(g);
    std::vector<uint64_t> shells(nv + 2);
    uint64_t ns = 0;
    uint32_t dg = 0;
    if (g2h_graph_kcore_recode(g, nullptr, shells.data(), &ns, &dg)) fail("kcore recode");
    const std::string bin = prefix_of(f) + ".bin";
    if (g2h_graph_write_bin(g, bin.c_str())) fail("cannot write " + bin);
    std::vector<uint32_t> bins(nv ? nv : 1);
    if (g2h_kcore_bins(shells.data(), ns, nb, bins.data(), nv)) fail("kcore bins");
    const std::string vb = bin + ".2.hop.vbmap";
    if (g2h_write_vbmap(vb.c_str(), nb, bins.data(), nv)) fail("cannot write " + vb);
    std::printf("degeneracy %u; wrote %s and %s\n", dg, bin.c_str(), vb.c_str());
    g2h_graph_free(g);
    return 0;
}

int kmeans(int argc, char** argv) {
    const std::string f = opt(argc, argv, "-f", "./data/com-dblp.ungraph.txt");
    const std::string ks = opt(argc, argv, "-k", "4");
    const uint32_t iters = (uint32_t)std::atoi(opt(argc, argv, "-i", "20").c_str());
    const uint64_t seed = std::strtoull(opt(argc, argv, "-seed", "0x6732A1D0C000000B").c_str(), nullptr, 0);
    const std::string out = opt(argc, argv, "-o", f + ".kmeans.") + ks;
    const size_t x = ks.find('x');
    const uint32_t k1 = (uint32_t)std::atoi(ks.substr(0, x).c_str());
    const uint32_t k2 = x == std::string::npos ? 1 : (uint32_t)std::atoi(ks.substr(x + 1).c_str());
    g2h_graph* g = load(f);
    const uint64_t nv = g2h_graph_nv(g);
    std::vector<uint32_t> cl(nv ? nv : 1);
    uint32_t n = 0;
    if (g2h_kmeans(g2h_graph_row_ptrs(g), nv, g2h_graph_cols(g), k1, k2, iters, seed, cl.data(), &n)) fail("kmeans");
    FILE* o = std::fopen(out.c_str(), "w");
    if (!o) fail("cannot write " + out);
    for (uint64_t v = 0; v < nv; v++) std::fprintf(o, "%u\n", cl[v]);
    std::fclose(o);
    std::printf("%u clusters; wrote %s\n", n, out.c_str());
    g2h_graph_free(g);
    return 0;
}

int view_pack(int argc, char** argv) {
    const std::string gf = opt(argc, argv, "-gf", "./data/com-dblp.ungraph.txt.bin");
    const std::string pf = opt(argc, argv, "-pf", "./data/com-dblp.ungraph.txt.bin.kmeans.4");
    const int hop = std::atoi(opt(argc, argv, "-h", "2").c_str());
    const double mem = std::atof(opt(argc, argv, "-m", "10").c_str());
    const uint32_t pool = (uint32_t)std::atoi(opt(argc, argv, "-s", "1").c_str());
    const int by_degree = std::atoi(opt(argc, argv, "-o", "0").c_str());
    const int dump = std::atoi(opt(argc, argv, "-d", "1").c_str());
    g2h_graph* g = load(gf);
    const uint64_t nv = g2h_graph_nv(g);
    const uint64_t* rp = g2h_graph_row_ptrs(g);
    std::vector<uint32_t> cl(nv ? nv : 1);
    uint32_t ncl = 0;
    {
        std::ifstream in(pf);
        if (!in) fail("cannot read " + pf);
        for (uint64_t v = 0; v < nv; v++) {
            long long c = -1;
            if (!(in >> c) || c < 0 || c > 65535) fail("partition file " + pf + " is short or holds a bad id");
            cl[v] = (uint32_t)c;
            ncl = std::max<uint32_t>(ncl, (uint32_t)c + 1);
        }
    }
    const uint64_t max_entries = (uint64_t)(mem * 1000 * 1000 * 1000) / 4;
    std::vector<uint32_t> bins(nv ? nv : 1);
    uint32_t nb = 0;
    if (g2h_view_pack(rp, nv, g2h_graph_cols(g), cl.data(), ncl, hop, max_entries, pool, by_degree, bins.data(), &nb))
        fail("view packing");
    std::printf("total %u bins after merging\n", nb);
    if (dump) {
        const std::string base = gf + "." + std::to_string(hop) + ".hop";
        std::vector<uint64_t> sources(nb, 0);
        for (uint64_t v = 0; v < nv; v++) sources[bins[v]]++;
        FILE* st = std::fopen((base + ".vbstat").c_str(), "w");
        if (!st) fail("cannot write " + base + ".vbstat");
        std::fprintf(st, "view bin count: %u\n", nb);
        for (uint32_t b = 0; b < nb; b++) std::fprintf(st, "view bin id: %u, sources: %llu\n", b, (unsigned long long)sources[b]);
        std::fclose(st);
        if (g2h_write_vbmap_lines((base + ".vbmap").c_str(), nb, bins.data(), nv)) fail("cannot write " + base + ".vbmap");
        std::printf("wrote %s.vbmap\n", base.c_str());
    }
    g2h_graph_free(g);
    return 0;
}

// converter/toGAMMA.cpp: <prefix>.col (u32 V, u64 row_ptrs[V+1]),
// <prefix>.dst (u64 E, u32 cols[E]), <prefix>.vlabel (u32 V, V bytes of label 1)
int to_gamma(const std::string& f) {
    g2h_graph* g = load(f);
    const std::string pre = prefix_of(f);
    const uint64_t nv = g2h_graph_nv(g), ne = g2h_graph_ne(g);
    const uint32_t v32 = (uint32_t)nv;
    FILE* o = std::fopen((pre + ".col").c_str(), "wb");
    if (!o) fail("cannot write " + pre + ".col");
    std::fwrite(&v32, 4, 1, o);
    std::fwrite(g2h_graph_row_ptrs(g), 8, nv + 1, o);
    std::fclose(o);
    o = std::fopen((pre + ".dst").c_str(), "wb");
    if (!o) fail("cannot write " + pre + ".dst");
    std::fwrite(&ne, 8, 1, o);
    if (ne) std::fwrite(g2h_graph_cols(g), 4, ne, o);
    std::fclose(o);
    o = std::fopen((pre + ".vlabel").c_str(), "wb");
    if (!o) fail("cannot write " + pre + ".vlabel");
    std::fwrite(&v32, 4, 1, o);
    const std::vector<uint8_t> label(nv, 1);
    if (nv) std::fwrite(label.data(), 1, nv, o);
    std::fclose(o);
    g2h_graph_free(g);
    return 0;
}

// converter/toGSI.cpp: <prefix>.g, "t # 0", "V E 1 1", a "v i 1" line per
// vertex, an "e i j 1" line per edge (i < j), "t # -1"
// converter/toPangolin.cpp: <prefix>.pangolin.txt, "i+1 j+1" per edge (i < j)
int to_text(const std::string& f, bool gsi) {
    g2h_graph* g = load(f);
    const std::string out = prefix_of(f) + (gsi ? ".g" : ".pangolin.txt");
    const uint64_t nv = g2h_graph_nv(g), ne = g2h_graph_ne(g);
    const uint64_t* rp = g2h_graph_row_ptrs(g);
    const uint32_t* cols = g2h_graph_cols(g);
    FILE* o = std::fopen(out.c_str(), "w");
    if (!o) fail("cannot write " + out);
    if (gsi) {
        std::fprintf(o, "t # 0\n%llu %llu 1 1\n", (unsigned long long)nv, (unsigned long long)(ne / 2));
        for (uint64_t i = 0; i < nv; i++) std::fprintf(o, "v %llu 1\n", (unsigned long long)i);
    }
    for (uint64_t i = 0; i < nv; i++)
        for (uint64_t j = rp[i]; j < rp[i + 1]; j++)
            if (i < cols[j]) {
                if (gsi) std::fprintf(o, "e %llu %u 1\n", (unsigned long long)i, cols[j]);
                else std::fprintf(o, "%llu %llu\n", (unsigned long long)i + 1, (unsigned long long)cols[j] + 1);
            }
    if (gsi) std::fprintf(o, "t # -1\n");
    std::fclose(o);
    g2h_graph_free(g);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::string name = argv[0];
    name = name.substr(name.rfind('/') + 1);
    if (name == "preprocess") return preprocess(argc, argv);
    if (name == "kcore") return kcore(argc, argv);
    if (name == "kmeans") return kmeans(argc, argv);
    if (name == "view_pack") return view_pack(argc, argv);
    if (name == "toGAMMA" || name == "toGSI" || name == "toPangolin") {
        if (argc < 2) {
            std::printf("usage: %s <graph file>\n", name.c_str());
            return -1;
        }
        return name == "toGAMMA" ? to_gamma(argv[1]) : to_text(argv[1], name == "toGSI");
    }
    std::printf("call this program as preprocess, kcore, kmeans, view_pack, toGAMMA, toGSI or toPangolin\n");
    return -1;
}
