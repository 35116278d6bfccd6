#pragma once
#define MVT int
#define MVT_Init_Value -1
