// Command-line driver with the reference's contract (tc/bs_tc/main.cu:6-20,
// tc/script/run_bs_tc.sh:3): bs_tc <graph_file> <device_id>. Files ending in .mtx are read as
// Matrix Market, .txt as text pairs, anything else as binary int32 pairs.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "graph.h"
#include "rtg_compat.h"

static bool ends_with(const std::string &s, const char *suf) {
    const std::string t(suf);
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

int main(int argc, char *argv[]) {
    if (argc != 3) {
        std::printf("Usage: ./bs_tc graph device_id\n");
        std::exit(1);
    }
    std::string graph_file = argv[1];
    const int device_id = std::atoi(argv[2]);
    Graph graph(device_id);
    if (ends_with(graph_file, ".mtx")) graph.ReadMtx(graph_file);
    else if (ends_with(graph_file, ".txt")) graph.ReadText(graph_file);
    else graph.ReadBinaryEdges(graph_file);
    graph.TCPreprocessing();
    graph.ConvertToCSR();
    CountOnGPU(graph.adjs_.data(), graph.offsets_.data(), graph.graph_info_.nodes, graph.graph_info_.edges, device_id);
    return 0;
}
