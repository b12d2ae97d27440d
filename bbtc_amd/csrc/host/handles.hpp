// Definitions of the opaque host handles shared by the host and device parts
// of the library.
#pragma once

#include "graph.hpp"

struct bbtc_hgraph {
  bbtc::Csr g;
};

struct bbtc_layout {
  bbtc::Layout L;
};
