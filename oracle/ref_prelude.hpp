// Test infrastructure only: force-included ahead of the reference's own sources
// when oracle/Makefile compiles them into oracle/_ref/.  It pulls in every
// standard header first and then opens class access, so the harness can read
// and write the model's private state (LSDCatchmentModel.hpp:427-816) between
// method calls.  Nothing in the product links this.
#ifndef HC_REF_PRELUDE_HPP
#define HC_REF_PRELUDE_HPP
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <limits>
#include <list>
#include <map>
#include <memory>
#include <numeric>
#include <queue>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include <omp.h>
#include <sys/stat.h>
#define private public
#define protected public
#endif
