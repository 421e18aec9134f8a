// Pass kernel variants, part 6 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): tuning alternatives for the large power-of-two passes.
// One translation unit per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT_T(64, 64, 128, 2),
    GL_VARIANT_T(64, 64, 64, 1),
    GL_VARIANT_T(32, 32, 128, 4),
    GL_VARIANT_T(16, 16, 512, 32),
};

const PassVariant *pass_variants_part6(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
