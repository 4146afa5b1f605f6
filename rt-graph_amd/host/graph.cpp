#include "graph.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace {
[[noreturn]] void die(const char *what) {
    std::cerr << "Error: " << what << ": " << rtg_last_error() << std::endl;
    std::exit(2);
}
}  // namespace

Graph::Graph(int device_id) : device_id_(device_id) {}
Graph::Graph() : device_id_(0) {}

Graph::~Graph() {
    if (g_) rtg_graph_destroy(g_);
    if (ctx_) rtg_destroy(ctx_);
}

void Graph::Ensure() {
    if (!ctx_ && rtg_create(device_id_, &ctx_) != RTG_OK) die("rtg_create");
}

void Graph::ReadBinaryEdges(std::string &input_file_path) {
    int64_t n = 0;
    if (rtg_read_binary_pairs(input_file_path.c_str(), nullptr, &n) != RTG_OK) die("ReadBinaryEdges");
    std::cout << "The size of graph file: " << n * 8 << " bytes\n";
    edge_list_.resize((size_t)n * 2);
    if (rtg_read_binary_pairs(input_file_path.c_str(), edge_list_.data(), &n) != RTG_OK) die("ReadBinaryEdges");
    graph_info_.edges = (int)n;
}

void Graph::ReadMtx(const std::string &input_file_path) {
    int64_t n = 0;
    if (rtg_read_mtx_pairs(input_file_path.c_str(), nullptr, &n) != RTG_OK) die("ReadMtx");
    edge_list_.resize((size_t)n * 2);
    if (rtg_read_mtx_pairs(input_file_path.c_str(), edge_list_.data(), &n) != RTG_OK) die("ReadMtx");
    graph_info_.edges = (int)n;
}

void Graph::ReadText(const std::string &input_file_path) {
    int64_t n = 0;
    if (rtg_read_text_pairs(input_file_path.c_str(), nullptr, &n) != RTG_OK) die("ReadText");
    edge_list_.resize((size_t)n * 2);
    if (rtg_read_text_pairs(input_file_path.c_str(), edge_list_.data(), &n) != RTG_OK) die("ReadText");
    graph_info_.edges = (int)n;
}

void Graph::TCPreprocessing() {
    Ensure();
    if (g_) {
        rtg_graph_destroy(g_);
        g_ = nullptr;
    }
    rtg_build_stats st{};
    if (rtg_graph_from_pairs(ctx_, edge_list_.data(), (int64_t)edge_list_.size() / 2,
                             RTG_PAIRS_RELABEL | RTG_PAIRS_ORIENT, &g_, &st) != RTG_OK)
        die("TCPreprocessing");
    std::cout << "Graph Info: nodes = " << st.nodes_in << ", edges = " << st.raw_pairs << std::endl;
    std::cout << "TC Preprocessing Time = " << st.t_total_ms << " ms\n";
}

void Graph::ConvertToCSR() {
    if (!g_) die("ConvertToCSR before TCPreprocessing");
    auto t0 = std::chrono::steady_clock::now();
    int32_t n = 0;
    int64_t m = 0;
    rtg_graph_info(g_, &n, &m, nullptr);
    offsets_.resize((size_t)n + 1);
    adjs_.resize((size_t)m);
    if (rtg_graph_download(g_, offsets_.data(), adjs_.data()) != RTG_OK) die("ConvertToCSR");
    graph_info_.nodes = n;
    graph_info_.edges = (int)m;
    std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t0;
    std::cout << "CSR converting time = " << dt.count() << " ms\n";
}

void Graph::Txt2Bin(const std::string &txt_file, const std::string &bin_file) {
    int64_t n = 0;
    if (rtg_read_text_pairs(txt_file.c_str(), nullptr, &n) != RTG_OK) die("Txt2Bin");
    std::vector<int32_t> p((size_t)n * 2);
    if (rtg_read_text_pairs(txt_file.c_str(), p.data(), &n) != RTG_OK) die("Txt2Bin");
    if (rtg_write_binary_pairs(bin_file.c_str(), p.data(), n) != RTG_OK) die("Txt2Bin");
    std::cout << "Edge list has been successfully converted to binary." << std::endl;
}
