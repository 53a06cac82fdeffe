// app.cpp -- the ten applications {bfs,cc,sssp,sswp,pr}-{sync,async}
// (reference: subway/*.cu).  APP_NAME is set per binary by the Makefile.
// Output lines follow the reference: read time, node/edge counts,
// "Processing finished in <s> (s).", "Number of iterations = <n>", the
// first 30 results, and --output saving.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <iostream>
#include <string>

#include "subway/argument_parsing.hpp"
#include "subway/engine.hpp"
#include "subway/gpu_error_check.hpp"
#include "subway/graph.hpp"
#include "subway/subway_utilities.hpp"

#ifndef APP_NAME
#error "APP_NAME must name the application, e.g. \"sssp-async\""
#endif

using namespace subway;

static double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

static void report(const RunStats &s, int mode)
{
    std::cout << "Processing finished in " << s.msTotal / 1000.0 << " (s).\n";
    if (mode == SYNC) std::cout << "Number of iterations = " << s.globalIterations << std::endl;
    else std::cout << "Global iterations = " << s.globalIterations << ", inner rounds = " << s.innerRounds << std::endl;
    std::cout << "H2D edge bytes = " << s.h2dBytes << " (DMA " << s.h2dDmaBytes << ", GPU gather "
              << s.h2dGatherBytes << "), partitions = " << s.partitions << std::endl;
}

template <class G, class Run>
static int run_app(ArgumentParser &args, G &graph, Run run, RunOptions &o)
{
    auto t0 = std::chrono::steady_clock::now();
    graph.ReadGraph();
    std::cout << "Graph Reading finished in " << seconds_since(t0) << " (s).\n";
    RunStats s = run(graph, o);
    report(s, o.mode);
    utilities::PrintResults(graph.value, std::min(30u, graph.num_nodes));
    if (args.hasOutput) utilities::SaveResults(args.output, graph.value, graph.num_nodes);
    return 0;
}

int main(int argc, char **argv)
{
    const std::string name = APP_NAME;
    const size_t dash = name.find('-');
    const std::string an = name.substr(0, dash), mn = name.substr(dash + 1);
    int algo = an == "bfs" ? BFS : an == "cc" ? CC : an == "sssp" ? SSSP : an == "sswp" ? SSWP : PR;
    int mode = mn == "async" ? ASYNC : SYNC;
    // cc-sync takes no --source (subway/cc-sync.cu:17)
    ArgumentParser args(argc, argv, !(algo == CC && mode == SYNC), false);
    RunOptions o;
    o.algo = algo;
    o.mode = mode;
    o.source = (uint)args.sourceNode;
    o.budgetBytes = args.budgetBytes;
    o.numBuffers = args.numBuffers;
    o.acc = args.threshold;
    o.transfer = args.transfer;
    o.policy = args.policy;
    o.verbose = args.verbose;
    try {
        if (algo == PR) {
            GraphPR<OutEdge> g(args.input, true);
            return run_app(args, g, [](GraphPR<OutEdge> &gr, RunOptions &op) { return runPageRank(gr, op); }, o);
        }
        if (algo == SSSP || algo == SSWP) {
            Graph<OutEdgeWeighted> g(args.input, true);
            return run_app(args, g, [](Graph<OutEdgeWeighted> &gr, RunOptions &op) { return runLabel(gr, op); }, o);
        }
        Graph<OutEdge> g(args.input, false);
        return run_app(args, g, [](Graph<OutEdge> &gr, RunOptions &op) { return runLabel(gr, op); }, o);
    } catch (const Error &e) {
        std::cerr << e.what() << std::endl;
        return e.code > 0 ? e.code : 255;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 255;
    }
}
