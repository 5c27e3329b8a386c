// kate_graph.hpp -- weighted digraph + shortest path used by Kate's passes.
//
// Kate::Graph keeps the reference's dense interface
// (backendagnosticcode/graphUtil.hpp:11-80).  The passes use SparsePath, which
// runs the same Dijkstra over adjacency lists: nodes are expanded in
// (distance, index) order, a node's predecessor is the first node that
// reaches it and is replaced only by a strictly shorter path, and the search
// stops when the target is popped.  Given edges listed in ascending target
// order this visits nodes and breaks ties exactly as the dense scan does,
// so the fused/grouped programs are the reference's.
#ifndef KATE_AMD_GRAPH_HPP
#define KATE_AMD_GRAPH_HPP

#include <functional>
#include <iostream>
#include <queue>
#include <utility>
#include <vector>

namespace Kate {

class SparsePath {
public:
    explicit SparsePath(int n) : adj_((size_t)n) {}
    int size() const { return (int)adj_.size(); }
    // Edges must be added in ascending `to` order per source (the passes do).
    void edge(int from, int to, double w) { adj_[(size_t)from].emplace_back(to, w); }
    // Node sequence from src to dst, or {-1} when dst is unreachable.
    std::vector<int> path(int src, int dst) const
    {
        const int n = size();
        typedef std::pair<double, int> Entry;
        std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> open;
        std::vector<double> dist((size_t)n, 0.0);
        std::vector<int> pred((size_t)n, -1);
        pred[(size_t)src] = -2;
        open.push(Entry(0.0, src));
        while (!open.empty()) {
            int u = open.top().second;
            open.pop();
            if (u == dst) break;
            for (const auto& e : adj_[(size_t)u]) {
                const int k = e.first;
                const double w = e.second;
                if (w < 0) continue;
                if (pred[(size_t)k] == -1 || dist[(size_t)k] > dist[(size_t)u] + w) {
                    pred[(size_t)k] = u;
                    dist[(size_t)k] = dist[(size_t)u] + w;
                    open.push(Entry(dist[(size_t)k], k));
                }
            }
        }
        if (pred[(size_t)dst] == -1) return {-1};
        std::vector<int> out;
        for (int v = dst; v != -2; v = pred[(size_t)v]) out.push_back(v);
        return std::vector<int>(out.rbegin(), out.rend());
    }

private:
    std::vector<std::vector<std::pair<int, double>>> adj_;
};

// Dense adjacency with -1 for "no edge" (graphUtil.hpp:11-37).
class Graph {
public:
    int n;
    double* weight;
    explicit Graph(int nodes) : n(nodes), store_((size_t)nodes * nodes, -1.0) { weight = store_.data(); }
    Graph(const Graph& o) : n(o.n), store_(o.store_) { weight = store_.data(); }
    Graph& operator=(const Graph& o)
    {
        n = o.n;
        store_ = o.store_;
        weight = store_.data();
        return *this;
    }
    double operator()(int i, int j) const { return weight[(size_t)i * n + j]; }
    void operator()(int i, int j, double a) { weight[(size_t)i * n + j] = a; }
    void print() const
    {
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) std::cout << weight[(size_t)i * n + j] << " ";
            std::cout << std::endl;
        }
        std::cout << std::endl;
    }
    std::vector<int> pathFinding(int i, int j) const
    {
        SparsePath sp(n);
        for (int u = 0; u < n; u++)
            for (int k = 0; k < n; k++) {
                double w = weight[(size_t)u * n + k];
                if (w >= 0) sp.edge(u, k, w);
            }
        return sp.path(i, j);
    }

private:
    std::vector<double> store_;
};

}  // namespace Kate

#endif
