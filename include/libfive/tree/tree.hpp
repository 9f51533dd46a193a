// libfive::Tree, as declared by libfive/include/libfive/tree/tree.hpp
// (refcounted expression nodes, the same objects as the C API's
// libfive_tree handles).  The declaration lives with the library sources.
#pragma once

#include "../../../libfive_amd/csrc/tree.hpp"
