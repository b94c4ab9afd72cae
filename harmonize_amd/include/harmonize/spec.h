// Reads a program specification's optional members, falling back to defaults
// when a member is absent. Same contract as the reference's MEMBER_SWITCH /
// CONST_SWITCH (program/async.h:26-36, 101-131; program/event.h:22-32,
// 59-75): OpSet, DeviceState, GroupState, ThreadState, AdrType and the
// *_SIZE constants are all optional. A state declared as a member template
// (as the reference's barrier demo does) is not a type and falls back to
// VoidState, as in the reference.
#pragma once

#include "promise.h"

namespace hz {
namespace spec {

#define HZ_TYPE_OR(NAME)                                                     \
    template <typename S, typename D, typename = void>                       \
    struct NAME##Or {                                                        \
        using type = D;                                                      \
    };                                                                       \
    template <typename S, typename D>                                        \
    struct NAME##Or<S, D, std::void_t<typename S::NAME>> {                   \
        using type = typename S::NAME;                                       \
    };

HZ_TYPE_OR(OpSet)
HZ_TYPE_OR(DeviceState)
HZ_TYPE_OR(GroupState)
HZ_TYPE_OR(ThreadState)
HZ_TYPE_OR(AdrType)
#undef HZ_TYPE_OR

#define HZ_CONST_OR(NAME)                                                    \
    template <typename S, size_t D, typename = void>                         \
    struct NAME##Or {                                                        \
        static constexpr size_t value = D;                                   \
        static constexpr bool given = false;                                 \
    };                                                                       \
    template <typename S, size_t D>                                          \
    struct NAME##Or<S, D, std::void_t<decltype(S::NAME)>> {                  \
        static constexpr size_t value = (size_t)S::NAME;                     \
        static constexpr bool given = true;                                  \
    };

HZ_CONST_OR(STASH_SIZE)
HZ_CONST_OR(FRAME_SIZE)
HZ_CONST_OR(POOL_SIZE)
HZ_CONST_OR(STACK_SIZE)
HZ_CONST_OR(GROUP_SIZE)
// Groups (wavefronts) launched per workgroup by the async program type; they
// share the workgroup's __shared__ memory. Not a reference constant.
HZ_CONST_OR(GROUPS_PER_BLOCK)
// Nonzero: a group that holds several links hands some to groups that have
// none (for work that grows from few seeds). Not a reference constant.
HZ_CONST_OR(SHARE_WORK)
// Nonzero: a group runs its ready links oldest first (the order in which
// they filled), rather than by stash position. Work whose later promises can
// supersede earlier ones (a BFS level lowered again) wastes less. Not a
// reference constant.
HZ_CONST_OR(OLDEST_FIRST)
// Nonzero, with STACK_SIZE depth levels: run and pull the shallowest level
// first instead of the deepest. Not a reference constant.
HZ_CONST_OR(SHALLOW_FIRST)
// Zero: the async program type never forwards links (runs a link's
// successor promises straight from registers); forwarding is on by default
// for flat programs. Not a reference constant.
HZ_CONST_OR(FORWARD_LINKS)
#undef HZ_CONST_OR

// The operation set a program type executes: the specification's OpSet, plus
// ReturnOp when an operation returns a value (hz::ExecOpSet).
template <typename S>
using ProgramOpSet = typename ExecOpSet<typename OpSetOr<S, OpUnion<>>::type>::type;

}  // namespace spec
}  // namespace hz
