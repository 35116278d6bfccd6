// BFS message/level type: a level is an int, -1 before the vertex is reached.
#pragma once
#define MVT int
#define MVT_Init_Value -1
