// Combiner check with float messages.  Usage: combine_float_app <graph> <out> <add|min|max|mul>
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"
#include "../common/combine_main.h"

int main(int argc, char** argv) { return CombineMain(argc, argv, MEDUSA_COMBINE_FLOAT, "combine_float_app"); }
