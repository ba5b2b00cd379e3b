// Umbrella header (include/GRE.hpp:1-17 in the GRE tree).
#ifndef GRE_AMD_GRE_HPP
#define GRE_AMD_GRE_HPP
#include "Engine/Engine.hpp"
#include "VertexCode/VertexCode.hpp"
#include "Graph/DistributedGraph.hpp"
#include "Graph/VertexSet.hpp"
#include "Comm/CommUnderlying.hpp"
#include "Comm/BufferingComm.hpp"
#include "Types.hpp"
#include "Status.hpp"
#endif
