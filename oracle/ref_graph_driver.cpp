// ref_graph_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own dynamic-graph ingest (Graph_dynamic.h, compiled
// from the reference tree by oracle/Makefile) over a list of snapshot edge
// files and dumps, after every snapshot, the degree, former degree and row
// contents of every vertex.  Its output is what the committed ingest golden
// fixtures under tests/golden/ were made from.
//
// usage: ref_graph u|d <n> <snapshot file>...
// output lines:  S <snapshot index> <m>
//                V <dir> <vertex> <degree> <former degree> : <row...>
// The reference includes <algorithm> before this header in every translation
// unit (e.g. Tree_u_NC.cpp:11,18); the header itself relies on it.
#include <algorithm>
#include "Graph_dynamic.h"

static void dump_rows(int dir, int n, int *deg, int *former, int **rows)
{
    for (int v = 0; v < n; v++) {
        printf("V %d %d %d %d :", dir, v, deg[v], former[v]);
        for (int j = 0; j < deg[v]; j++) printf(" %d", rows[v][j]);
        printf("\n");
    }
}

int main(int argc, char **argv)
{
    if (argc < 4) { fprintf(stderr, "usage: %s u|d n files...\n", argv[0]); return 2; }
    bool directed = argv[1][0] == 'd';
    int n = atoi(argv[2]);
    std::cout.setstate(std::ios_base::failbit);   // silence the reference's progress prints
    if (directed) {
        Graph *g = new Graph();
        g->initializeDirectedDynamicGraph(n);
        // The reference compares these row pointers to NULL without
        // initialising them (Graph_dynamic.h:100-104); start them empty.
        for (int i = 0; i < n; i++) { g->inAdjList[i] = NULL; g->outAdjList[i] = NULL; }
        for (int f = 3; f < argc; f++) {
            std::vector<std::pair<int, int>> ev;
            g->inputDirectedDynamicGraph(argv[f], ev);
            printf("S %d %d\n", f - 3, g->m);
            dump_rows(0, n, g->outdegree, g->former_outdegree, g->outAdjList);
            dump_rows(1, n, g->indegree, g->former_indegree, g->inAdjList);
        }
    } else {
        UGraph *g = new UGraph();
        g->initializeDynamicUGraph(n);
        for (int i = 0; i < n; i++) g->AdjList[i] = NULL;
        for (int f = 3; f < argc; f++) {
            std::vector<std::pair<int, int>> ev;
            g->inputDynamicGraph(argv[f], ev);
            printf("S %d %d\n", f - 3, g->m);
            dump_rows(0, n, g->degree, g->former_degree, g->AdjList);
        }
    }
    return 0;
}
