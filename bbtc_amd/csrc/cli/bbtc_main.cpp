#include "bbtc.h"
int main() { return 0; }
