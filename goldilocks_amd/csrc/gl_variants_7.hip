// Pass kernel variants, part 7 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): three-stage passes; the last two are 1024 points in
// 256-thread workgroups of two instances, twice the workgroups of the
// 512-thread form for cache-resident lengths.  One translation unit per
// part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT3(16, 16, 16, 1024, 4),
    GL_VARIANT3(16, 16, 16, 512, 2),
    GL_VARIANT3(8, 8, 8, 512, 8),
    GL_VARIANT3(16, 16, 8, 512, 4),
    GL_VARIANT3(16, 8, 8, 512, 4),
    GL_VARIANT3(16, 8, 8, 256, 2),
    GL_VARIANT3(8, 8, 16, 256, 2),
};

const PassVariant *pass_variants_part7(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
