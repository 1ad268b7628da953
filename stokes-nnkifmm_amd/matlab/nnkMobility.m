function id = nnkMobility(swimmers, B, BT, A1, A2, precon)
%NNKMOBILITY The library's mobility operator [S BT; B 0] of
%   (swimmers.FMM, B, BT) -- with the coronaRes-4 operator of the
%   preconditioner when precon is true -- built on first use and reused
%   while swimmers.FMM, B and BT are unchanged (the reference forms them
%   once per mobilityProblem call, Mobility/mobilityProblem.m:31-35).
persistent cache
fmm = swimmers.FMM;
fid = fmm.handle(fmm.arguments.coronaRes);
f4 = uint64([]);
if precon
    f4 = fmm.handle(4);
end
if ~isempty(cache) && cache.fid == fid && isequal(cache.f4, f4) && isequal(cache.B, B) && isequal(cache.BT, BT)
    id = cache.h.id;
    return
end
if isempty(A1)
    A1 = B * B.';
end
if isempty(A2)
    A2 = BT.' * BT;
end
h = nnkHandle(nnkifmm_mex('mobility', fid, f4, sparse(B.'), sparse(BT.'), full(A1).', full(A2).'));
cache = struct('fid', fid, 'f4', f4, 'B', B, 'BT', BT, 'h', h);
id = h.id;
end
