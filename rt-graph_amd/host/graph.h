// Host-side graph front end with the reference's interface (tc/graph/include/graph.h:10-30,
// graph_info_t from tc/utils/model.h:10-13): read an edge list, preprocess it into the
// degree-oriented CSR, expose adjs_/offsets_/graph_info_. The preprocessing runs on the GPU
// through librtg; ReadMtx and ReadText add the .mtx and text formats.
#pragma once

#include <string>
#include <vector>

#include "rtg.h"

typedef struct GraphInfo {
    int nodes;
    int edges;
} graph_info_t;

class Graph {
public:
    explicit Graph(int device_id);
    Graph();
    ~Graph();
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;

    void ReadBinaryEdges(std::string &input_file_path);
    void ReadMtx(const std::string &input_file_path);
    void ReadText(const std::string &input_file_path);
    void TCPreprocessing();
    void ConvertToCSR();
    //* edge list: text file to binary file
    static void Txt2Bin(const std::string &txt_file, const std::string &bin_file);

    std::vector<int> adjs_;
    std::vector<int> offsets_;
    graph_info_t graph_info_{0, 0};

private:
    void Ensure();
    std::vector<int32_t> edge_list_;  // flattened (src, dst) pairs as read
    int device_id_ = 0;
    rtg_ctx *ctx_ = nullptr;
    rtg_graph *g_ = nullptr;
};
